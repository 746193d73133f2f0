// Pair-force kernels over the device neighbour list (force.f90), replacing
// the README pair loop (README.md:86-103) with one launch sequence per
// potential.
//
//  * isotropic potentials and Gay-Berne: LPP lanes per particle walk the
//    particle's sorted full row; each lane evaluates the canonical pair
//    (min, max) exactly as the reference would and keeps the half it needs
//    (+FXIJ / TXI when i is the reference's I, -FXIJ / TXJ when it is J);
//    lanes combine with cross-lane shuffles in a fixed order.  No atomics on
//    forces: results are deterministic run to run.
//  * ECP: a conservative FP32 screen over the half list rejects pairs that
//    provably give zero (F(psi) > contact threshold => the Brent minimum lies
//    beyond contact), survivors are compacted and solved one lane per pair in
//    FP64 along the reference's Brent path, results land in a per-pair buffer
//    and a per-particle gather applies them.
#include "sx_ecp.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

struct Params {
    double p[8];
};

// ---------------------------------------------------------------------------
// isotropic / GB gather
// ---------------------------------------------------------------------------
template <int KIND, int LPP, bool TORQUE, typename T>
__global__ void __launch_bounds__(256) k_gather(DState s, const int *row_start, const int *nbr, Params P,
                                                double *pe_acc)
{
    int gid = blockIdx.x * blockDim.x + threadIdx.x;
    int i = gid / LPP, l = gid % LPP;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};  // fx fy fz tx ty tz pe
    if (i < s.n) {
        int b = row_start[i], e = row_start[i + 1];
        for (int q = b + l; q < e; q += LPP) {
            int j = nbr[q];
            PairOutT<T> o;
            if (j > i) {
                d_pair_t<KIND>(s, i, j, P.p, o);
                acc[0] = acc[0] + o.f[0]; acc[1] = acc[1] + o.f[1]; acc[2] = acc[2] + o.f[2];
                if (TORQUE) { acc[3] = acc[3] + o.ti[0]; acc[4] = acc[4] + o.ti[1]; acc[5] = acc[5] + o.ti[2]; }
                acc[6] = acc[6] + o.pe;
            } else {
                d_pair_t<KIND>(s, j, i, P.p, o);
                acc[0] = acc[0] - o.f[0]; acc[1] = acc[1] - o.f[1]; acc[2] = acc[2] - o.f[2];
                if (TORQUE) { acc[3] = acc[3] + o.tj[0]; acc[4] = acc[4] + o.tj[1]; acc[5] = acc[5] + o.tj[2]; }
            }
        }
    }
#pragma unroll
    for (int off = LPP / 2; off > 0; off >>= 1)
#pragma unroll
        for (int c = 0; c < 7; ++c) acc[c] = acc[c] + __shfl_xor(acc[c], off, LPP);
    if (l == 0 && i < s.n) {
        s.f[SX_FX][i] = s.f[SX_FX][i] + acc[0];
        s.f[SX_FY][i] = s.f[SX_FY][i] + acc[1];
        s.f[SX_FZ][i] = s.f[SX_FZ][i] + acc[2];
        if (TORQUE) {
            s.f[SX_TX][i] = s.f[SX_TX][i] + acc[3];
            s.f[SX_TY][i] = s.f[SX_TY][i] + acc[4];
            s.f[SX_TZ][i] = s.f[SX_TZ][i] + acc[5];
        }
        pe_acc[i] = pe_acc[i] + acc[6];
    }
}

// GbSelf records of every particle (d_gb_tab)
__global__ void k_gb_self(DState s, int mu, double *tab)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    GbSelf g;
    gb_self(s, i, mu, g);
    gb_self_store(g, tab + (long)GB_TAB * i);
}

template <int VARIANT, typename T>
__device__ __forceinline__ void gb_tab_t(const DState &s, int a, int b, const double *p, const double *tab,
                                         PairOutT<T> &o)
{
    if constexpr (sizeof(T) == sizeof(double)) {
        d_gb_tab<VARIANT>(s, a, b, p, tab, o);
    } else {
        static_assert(VARIANT == 0, "the FP32 pair mode covers gb_ only");
        d_gb_tab_f32(s, a, b, p, tab, o);
    }
}

// k_gather for the Gay-Berne family with the per-particle factors read from
// the table: the same pairs in the same order, the same values (T = float:
// gb_ in the FP32 pair mode's mixed form)
template <int VARIANT, int LPP, typename T>
__global__ void __launch_bounds__(256) k_gather_gb(DState s, const int *row_start, const int *nbr, Params P,
                                                   const double *tab, double *pe_acc)
{
    int gid = blockIdx.x * blockDim.x + threadIdx.x;
    int i = gid / LPP, l = gid % LPP;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};  // fx fy fz tx ty tz pe
    if (i < s.n) {
        int b = row_start[i], e = row_start[i + 1];
        for (int q = b + l; q < e; q += LPP) {
            int j = nbr[q];
            PairOutT<T> o;
            if (j > i) {
                gb_tab_t<VARIANT>(s, i, j, P.p, tab, o);
                acc[0] = acc[0] + o.f[0]; acc[1] = acc[1] + o.f[1]; acc[2] = acc[2] + o.f[2];
                acc[3] = acc[3] + o.ti[0]; acc[4] = acc[4] + o.ti[1]; acc[5] = acc[5] + o.ti[2];
                acc[6] = acc[6] + o.pe;
            } else {
                gb_tab_t<VARIANT>(s, j, i, P.p, tab, o);
                acc[0] = acc[0] - o.f[0]; acc[1] = acc[1] - o.f[1]; acc[2] = acc[2] - o.f[2];
                acc[3] = acc[3] + o.tj[0]; acc[4] = acc[4] + o.tj[1]; acc[5] = acc[5] + o.tj[2];
            }
        }
    }
#pragma unroll
    for (int off = LPP / 2; off > 0; off >>= 1)
#pragma unroll
        for (int c = 0; c < 7; ++c) acc[c] = acc[c] + __shfl_xor(acc[c], off, LPP);
    if (l == 0 && i < s.n) {
        s.f[SX_FX][i] = s.f[SX_FX][i] + acc[0];
        s.f[SX_FY][i] = s.f[SX_FY][i] + acc[1];
        s.f[SX_FZ][i] = s.f[SX_FZ][i] + acc[2];
        s.f[SX_TX][i] = s.f[SX_TX][i] + acc[3];
        s.f[SX_TY][i] = s.f[SX_TY][i] + acc[4];
        s.f[SX_TZ][i] = s.f[SX_TZ][i] + acc[5];
        pe_acc[i] = pe_acc[i] + acc[6];
    }
}

// two isotropic kinds over one list in one pass (one read of each row and its
// neighbours): each kind's sums are kept apart, reduced as k_gather reduces
// them and added in the order of the two loops, so forces and energies equal
// those of two k_gather launches with the same lanes per particle
template <int K1, int K2, int LPP, typename T>
__global__ void __launch_bounds__(256) k_gather2(DState s, const int *row_start, const int *nbr, Params P1,
                                                 Params P2, double *pe_acc)
{
    int gid = blockIdx.x * blockDim.x + threadIdx.x;
    int i = gid / LPP, l = gid % LPP;
    double a1[4] = {0, 0, 0, 0}, a2[4] = {0, 0, 0, 0};  // fx fy fz pe of each kind
    if (i < s.n) {
        int b = row_start[i], e = row_start[i + 1];
        for (int q = b + l; q < e; q += LPP) {
            int j = nbr[q];
            PairOutT<T> o1, o2;
            if (j > i) {
                d_pair_t<K1>(s, i, j, P1.p, o1);
                d_pair_t<K2>(s, i, j, P2.p, o2);
                a1[0] = a1[0] + o1.f[0]; a1[1] = a1[1] + o1.f[1]; a1[2] = a1[2] + o1.f[2];
                a1[3] = a1[3] + o1.pe;
                a2[0] = a2[0] + o2.f[0]; a2[1] = a2[1] + o2.f[1]; a2[2] = a2[2] + o2.f[2];
                a2[3] = a2[3] + o2.pe;
            } else {
                d_pair_t<K1>(s, j, i, P1.p, o1);
                d_pair_t<K2>(s, j, i, P2.p, o2);
                a1[0] = a1[0] - o1.f[0]; a1[1] = a1[1] - o1.f[1]; a1[2] = a1[2] - o1.f[2];
                a2[0] = a2[0] - o2.f[0]; a2[1] = a2[1] - o2.f[1]; a2[2] = a2[2] - o2.f[2];
            }
        }
    }
#pragma unroll
    for (int off = LPP / 2; off > 0; off >>= 1)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            a1[c] = a1[c] + __shfl_xor(a1[c], off, LPP);
            a2[c] = a2[c] + __shfl_xor(a2[c], off, LPP);
        }
    if (l == 0 && i < s.n) {
        s.f[SX_FX][i] = (s.f[SX_FX][i] + a1[0]) + a2[0];
        s.f[SX_FY][i] = (s.f[SX_FY][i] + a1[1]) + a2[1];
        s.f[SX_FZ][i] = (s.f[SX_FZ][i] + a1[2]) + a2[2];
        pe_acc[i] = (pe_acc[i] + a1[3]) + a2[3];
    }
}

// ---------------------------------------------------------------------------
// ECP
// ---------------------------------------------------------------------------
// per-particle records of every particle (ecp_pp_particle); thread 0 also
// clears the work counter of this call
__global__ void k_ecp_pp(DState s, double *pp, float *scr, int *nwork)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && nwork) *nwork = 0;
    if (i >= s.n) return;
    ecp_pp_particle(s, i, pp, scr);
}

struct EcpScreen {
    int variant;       // screening class: 0 cut (ecp, ecpw, ecpc, ecpwc), 1 ecpr
    double cutoff;     // reach of the cut forms: cutoff + wf
    double L[3], invL[3];
    float Lf[3], invLf[3];
    int periodic;
    int fvar;          // force variant: 0 ecp, 1 ecpr, 2 ecpw, 3 ecpc, 4 ecpwc
    EcpPar par;        // the potential's cutoff, weld width, chirality
};

static inline bool is_ecp_kind(int kind)
{
    return kind == SX_ECP || kind == SX_ECPR || kind == SX_ECPW || kind == SX_ECPC || kind == SX_ECPWC;
}

#define ECP_SCREEN_MARGIN 1e-2f

// Conservative screen of one half pair.  Brent only ever accepts a point
// with f(u) <= f(x) (force.f90:1132-1165), so its final lambda has F(lambda)
// >= F(psi); sigma = r / sqrt(F(lambda)) up to rounding, so F(psi) above the
// contact threshold proves a zero result.  F(psi) is evaluated in FP32 and a
// pair is kept unless F32(psi) exceeds the threshold by the 1 % margin, far
// above the FP32 error of a 3x3 solve whose condition number is bounded by
// (largest/smallest semi-axis)^2 (checked on the host).  Before the solve,
// F(psi) >= psi(1-psi) r^2 / lambda_max(G(psi)) with lambda_max(G) <=
// (1-psi) amax_a^2 + psi amax_b^2 rejects distant pairs without a solve.
// The separation is taken from the FP32 positions of the records: its error
// (a few 1e-6 in a box of edge 50) changes F by far less than the margin.
struct ScreenIn {
    float4 pa, pb;
    float4 A0, A1, B0, B1;
};

// the screen's inputs of pair (a, b): one record line each
__device__ __forceinline__ ScreenIn ecp_screen_load(const float *scr, int a, int b)
{
    const float4 *ra = reinterpret_cast<const float4 *>(scr + (long)a * ECP_SCR_STRIDE);
    const float4 *rb = reinterpret_cast<const float4 *>(scr + (long)b * ECP_SCR_STRIDE);
    return ScreenIn{ra[0], rb[0], ra[1], ra[2], rb[1], rb[2]};
}

__device__ __forceinline__ bool ecp_screen32(const ScreenIn &in, const EcpScreen &sc)
{
    float r0 = in.pa.x - in.pb.x, r1 = in.pa.y - in.pb.y, r2 = in.pa.z - in.pb.z;
    if (sc.periodic) {
        r0 = r0 - roundf(r0 * sc.invLf[0]) * sc.Lf[0];
        r1 = r1 - roundf(r1 * sc.invLf[1]) * sc.Lf[1];
        r2 = r2 - roundf(r2 * sc.invLf[2]) * sc.Lf[2];
    }
    float rsq = r0 * r0 + r1 * r1 + r2 * r2;
    float thr;
    if (sc.variant == 1) {
        thr = 1.0f;
    } else {
        float rmag = sqrtf(rsq);
        float den = rmag - (float)sc.cutoff + 1.0f;
        if (den <= 0.05f) return true;
        thr = (rmag / den) * (rmag / den);
    }
    thr = thr * (1.0f + ECP_SCREEN_MARGIN);
    const float psi = 0.381966011250105f;
    const float4 A0 = in.A0, A1 = in.A1, B0 = in.B0, B1 = in.B1;
    // aspect ratios beyond 30 are not screened (condition number of G)
    if (A1.z > 30.0f * A1.w || B1.z > 30.0f * B1.w) return true;
    float lmax = (1.0f - psi) * A1.z * A1.z + psi * B1.z * B1.z;
    if (psi * (1.0f - psi) * rsq > thr * lmax * (1.0f + ECP_SCREEN_MARGIN)) return false;
    float q = 1.0f - psi;
    float g11 = q * A0.x + psi * B0.x, g21 = q * A0.y + psi * B0.y, g31 = q * A0.z + psi * B0.z;
    float g22 = q * A0.w + psi * B0.w, g32 = q * A1.x + psi * B1.x, g33 = q * A1.y + psi * B1.y;
    float i11 = rsqrtf(g11);
    float l21 = g21 * i11, l31 = g31 * i11;
    float i22 = rsqrtf(g22 - l21 * l21);
    float l32 = (g32 - l31 * l21) * i22;
    float i33 = rsqrtf(g33 - l31 * l31 - l32 * l32);
    float y1 = r0 * i11;
    float y2 = (r1 - l21 * y1) * i22;
    float y3 = (r2 - l31 * y1 - l32 * y2) * i33;
    // r^T G^-1 r = |L^-1 r|^2
    float F = psi * (1.0f - psi) * (y1 * y1 + y2 * y2 + y3 * y3);
    return F <= thr;
}


// Compacts the segment's kept pairs in half-list order into the work list
__global__ void __launch_bounds__(256) k_ecp_screen(const float *scr, const int *half_i,
                                                    const int *half_j, const int *nnz_half, EcpScreen sc,
                                                    unsigned char *live, int *work, int *segcount)
{
    __shared__ int wcnt[4];
    long start, end;
    ecp_segment(*nnz_half, blockIdx.x, start, end);
    int t = threadIdx.x, lane = t & (SX_WAVE - 1), w = t / SX_WAVE;
    int run = 0;
    for (long base = start; base < end; base += 256) {
        long k = base + t;
        bool keep = false;
        if (k < end) {
            const int a = half_i[k], b = half_j[k];
            keep = ecp_screen32(ecp_screen_load(scr, a, b), sc);
            live[k] = keep ? 1 : 0;
        }
        unsigned long long m = __ballot(keep);
        if (lane == 0) wcnt[w] = __popcll(m);
        __syncthreads();
        int before = 0, total = 0;
        for (int q = 0; q < 4; ++q) {
            if (q < w) before += wcnt[q];
            total += wcnt[q];
        }
        if (keep) work[start + run + before + __popcll(m & ((1ull << lane) - 1ull))] = (int)k;
        run += total;
        __syncthreads();
    }
    if (t == 0) segcount[blockIdx.x] = run;
}

#define ECP_SOLVE_THREADS 64
#define ECP_SOLVE_WG 64    // workgroup of the solve stage (256, pairs spread over 4 wavefronts, measured 2x slower)

// one block per segment of the work list; stats (may be NULL) accumulates
// solved pairs, solver iterations, fallbacks and, for the learned solver,
// pairs whose warm start already proves them out of range (their live flag
// is cleared so the apply stage skips them)
template <int VARIANT>
__global__ void __launch_bounds__(ECP_SOLVE_WG) k_ecp_solve(DState s, const double *pp, long ld, const int *half_i,
                                                                 const int *half_j, const int *work, const int *segcount,
                                                                 const int *nnz_half, EcpScreen sc, int solver,
                                                                 const float *lam0, double *res, unsigned char *live,
                                                                 long *stats, EcpSlots sl)
{
    long start, end;
    ecp_segment(*nnz_half, blockIdx.x, start, end);
    long cnt = segcount[blockIdx.x];
    long my_iters = 0, my_solved = 0, my_fb = 0, my_skip = 0;
    // the segment's pairs dealt round-robin over the workgroup's wavefronts,
    // so each wavefront carries a few pairs and the CU interleaves them
    const int nw = blockDim.x / SX_WAVE, lane = threadIdx.x % SX_WAVE, wave = threadIdx.x / SX_WAVE;
    for (long u = (long)lane * nw + wave; u < cnt; u += blockDim.x) {
        long t = start + u;
        int k = work[t];
        int iters = 0, fb = 0;
        if (ecp_pair_solve<VARIANT>(s, pp, ld, k, half_i[k], half_j[k], solver, lam0 != nullptr,
                                    lam0 ? (double)lam0[t] : ECP_PSI, sc.par, res, live, sl, iters, fb)) {
            ++my_skip;
            continue;
        }
        my_iters += iters;
        my_solved += 1;
        my_fb += fb;
    }
    (void)end;
    if (!stats) return;
    // stats: [0] solved pairs, [1] iterations, [2] fallbacks, [3] proven zero by the warm start
    for (int off = SX_WAVE / 2; off > 0; off >>= 1) {
        my_iters += __shfl_xor(my_iters, off, SX_WAVE);
        my_solved += __shfl_xor(my_solved, off, SX_WAVE);
        my_fb += __shfl_xor(my_fb, off, SX_WAVE);
        my_skip += __shfl_xor(my_skip, off, SX_WAVE);
    }
    if ((threadIdx.x & (SX_WAVE - 1)) == 0 && (my_solved || my_skip)) {
        atomicAdd((unsigned long long *)&stats[0], (unsigned long long)my_solved);
        atomicAdd((unsigned long long *)&stats[1], (unsigned long long)my_iters);
        if (my_fb) atomicAdd((unsigned long long *)&stats[2], (unsigned long long)my_fb);
        if (my_skip) atomicAdd((unsigned long long *)&stats[3], (unsigned long long)my_skip);
    }
}

// per-particle application of the per-pair results, one thread per particle
// (ecp_particle_sums)
__global__ void __launch_bounds__(256) k_ecp_apply(DState s, EcpApplyArgs a, const unsigned char *live,
                                                   double *pe_acc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    // the particle's running sums, read before the pairs (independent of
    // them) so that their latency overlaps
    const double f0[7] = {s.f[SX_FX][i], s.f[SX_FY][i], s.f[SX_FZ][i], s.f[SX_TX][i],
                          s.f[SX_TY][i], s.f[SX_TZ][i], pe_acc[i]};
    double acc[7];
    ecp_particle_sums(i, a, live, acc);
    s.f[SX_FX][i] = f0[0] + acc[0];
    s.f[SX_FY][i] = f0[1] + acc[1];
    s.f[SX_FZ][i] = f0[2] + acc[2];
    s.f[SX_TX][i] = f0[3] + acc[3];
    s.f[SX_TY][i] = f0[4] + acc[4];
    s.f[SX_TZ][i] = f0[5] + acc[5];
    pe_acc[i] = f0[6] + acc[6];
}

__global__ void k_clear_slot_counts(int2 *slot, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) slot[(long)i * ECP_SLOT_LINE].x = 0;
}

// clears the slot counts that solves with no apply stage after them left
static void ecp_clear_slots(sx_ctx *c)
{
    hipLaunchKernelGGL(k_clear_slot_counts, dim3(sx_blocks(c->n, 256)), dim3(256), 0, c->stream, c->d_ecp_slot, c->n);
}

EcpApplyArgs sx_ecp_apply_args(sx_ctx *c)
{
    const NList &nl = c->nl;
    return EcpApplyArgs{c->d_ecp_slot, nl.row_start, nl.nbr, nl.lower, nl.half_start, c->d_pairbuf};
}

// explicit pairs (parity hook)
template <int KIND>
__global__ void k_pair_eval(DState s, long np, const int *pi, const int *pj, Params P, const double *pp, long ld,
                            int solver, double *out)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    int a = pi[t], b = pj[t];
    PairOut o;
    double lam = 0.0;
    constexpr int EV = KIND == SX_ECP ? 0 : KIND == SX_ECPR ? 1 : KIND == SX_ECPW ? 2 : KIND == SX_ECPC ? 3
                     : KIND == SX_ECPWC ? 4 : -1;
    if constexpr (EV >= 0) {
        double RIJ[3];
        pair_rij(s, a, b, RIJ);
        Sym3 A = load_sym(pp, ld, PP_AR, a), B = load_sym(pp, ld, PP_AR, b);
        double fpsi = ecp_optim_dist(ECP_PSI, A, B, RIJ);
        int it = 0, fb = 0;
        EcpPar ep;
        ep.cutoff = EV == 1 ? 0.0 : P.p[0];
        ep.wf = (EV == 2 || EV == 4) ? P.p[1] : 0.0;
        ep.chir = EV == 3 ? P.p[1] : (EV == 4 ? P.p[2] : 0.0);
        lam = ecp_solve_lambda<EV>(solver, pp, ld, a, b, RIJ, fpsi, ECP_PSI, it, fb);
        ecp_forces<EV>(s, pp, ld, a, b, ep, RIJ, lam, o);
    } else {
        for (int k = 0; k < 3; ++k) { o.ti[k] = 0.0; o.tj[k] = 0.0; }
        d_pair<KIND>(s, a, b, P.p, o);
    }
    double *r = out + t * 11;
    r[0] = o.pe;
    for (int k = 0; k < 3; ++k) { r[1 + k] = o.f[k]; r[4 + k] = o.ti[k]; r[7 + k] = o.tj[k]; }
    r[10] = lam;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static Params make_params(sx_ctx *c, int kind, const double *p, int np)
{
    Params P;
    for (int k = 0; k < 8; ++k) P.p[k] = 0.0;
    for (int k = 0; k < np && k < 8; ++k) P.p[k] = p[k];
    if (kind == SX_GB || kind == SX_GBR || kind == SX_GBW || kind == SX_GBC || kind == SX_GBWC) {
        if (np < 2) P.p[1] = 1.0;   // mu (force.f90:240)
        if (np < 3) P.p[2] = 2.0;   // nu
    }
    if (kind == SX_MORSE) P.p[4] = std::exp(-P.p[1] * (P.p[3] - P.p[2]));   // exp_cut (force.f90:224)
    (void)c;
    return P;
}

static void check_params(int kind, int np)
{
    int need = 0;
    switch (kind) {
    case SX_LJ: case SX_LJ12: case SX_DH: need = 3; break;
    case SX_LJW: case SX_MORSE: need = 4; break;
    case SX_WCA: need = 2; break;
    case SX_GB: case SX_ECP: need = 1; break;
    case SX_GBR: case SX_ECPR: need = 0; break;
    case SX_GBW: case SX_GBC: need = 4; break;   // cutoff, mu, nu, wf | chirality
    case SX_GBWC: need = 5; break;
    case SX_ECPW: case SX_ECPC: need = 2; break;
    case SX_ECPWC: need = 3; break;
    case SX_DIPOLE: case SX_COULOMB: need = 2; break;   // cutoff, dielectric
    default: throw sx_error{SX_ERR_ARG, "unknown force kind"};
    }
    if (np < need) throw sx_error{SX_ERR_ARG, "too few parameters for this force kind"};
}

template <int KIND, bool TORQUE, typename T>
static void launch_gather_t(sx_ctx *c, const Params &P, int lpp)
{
    int n = c->n;
    long threads = (long)n * lpp;
    int nb = sx_blocks(threads, 256);
    sx_ensure_part(c, nb, 1);
    DState s = c->dstate();
    switch (lpp) {
    case 1: hipLaunchKernelGGL((k_gather<KIND, 1, TORQUE, T>), dim3(nb), dim3(256), 0, c->stream, s, c->nl.row_start, c->nl.nbr, P, c->d_pe_acc); break;
    case 4: hipLaunchKernelGGL((k_gather<KIND, 4, TORQUE, T>), dim3(nb), dim3(256), 0, c->stream, s, c->nl.row_start, c->nl.nbr, P, c->d_pe_acc); break;
    case 8: hipLaunchKernelGGL((k_gather<KIND, 8, TORQUE, T>), dim3(nb), dim3(256), 0, c->stream, s, c->nl.row_start, c->nl.nbr, P, c->d_pe_acc); break;
    default: hipLaunchKernelGGL((k_gather<KIND, 16, TORQUE, T>), dim3(nb), dim3(256), 0, c->stream, s, c->nl.row_start, c->nl.nbr, P, c->d_pe_acc); break;
    }
}

// FP32 pair terms where the context asks for them and the kind has them
template <int KIND, bool TORQUE>
static void launch_gather(sx_ctx *c, const Params &P, int lpp)
{
    if constexpr (sx_has_f32<KIND>()) {
        if (c->pair_f32) {
            launch_gather_t<KIND, TORQUE, float>(c, P, lpp);
            return;
        }
    }
    launch_gather_t<KIND, TORQUE, double>(c, P, lpp);
}

// the Gay-Berne family: the per-particle table, then the gather over it
template <int VARIANT, typename T>
static void launch_gather_gb_t(sx_ctx *c, const Params &P, int lpp)
{
    int n = c->n;
    if (c->gb_tab_cap < n) {
        if (c->d_gb_tab) HIPCHK(hipFree(c->d_gb_tab));
        HIPCHK(hipMalloc(&c->d_gb_tab, sizeof(double) * GB_TAB * (size_t)n));
        c->gb_tab_cap = n;
        c->have_graph = false;
    }
    DState s = c->dstate();
    const double *tab = c->d_gb_tab;
    hipLaunchKernelGGL(k_gb_self, dim3(sx_blocks(n, 256)), dim3(256), 0, c->stream, s, (int)P.p[1], c->d_gb_tab);
    int nb = sx_blocks((long)n * lpp, 256);
    switch (lpp) {
    case 1: hipLaunchKernelGGL((k_gather_gb<VARIANT, 1, T>), dim3(nb), dim3(256), 0, c->stream, s, c->nl.row_start, c->nl.nbr, P, tab, c->d_pe_acc); break;
    case 4: hipLaunchKernelGGL((k_gather_gb<VARIANT, 4, T>), dim3(nb), dim3(256), 0, c->stream, s, c->nl.row_start, c->nl.nbr, P, tab, c->d_pe_acc); break;
    case 8: hipLaunchKernelGGL((k_gather_gb<VARIANT, 8, T>), dim3(nb), dim3(256), 0, c->stream, s, c->nl.row_start, c->nl.nbr, P, tab, c->d_pe_acc); break;
    default: hipLaunchKernelGGL((k_gather_gb<VARIANT, 16, T>), dim3(nb), dim3(256), 0, c->stream, s, c->nl.row_start, c->nl.nbr, P, tab, c->d_pe_acc); break;
    }
}

// FP32 for gb_ (variant 0) where the context asks for it
template <int VARIANT>
static void launch_gather_gb(sx_ctx *c, const Params &P, int lpp)
{
    if constexpr (VARIANT == 0) {
        if (c->pair_f32) {
            launch_gather_gb_t<0, float>(c, P, lpp);
            return;
        }
    }
    launch_gather_gb_t<VARIANT, double>(c, P, lpp);
}

static int choose_lpp(sx_ctx *c, int kind)
{
    // enough lanes per particle to fill the chip for small systems; one lane
    // per particle keeps the per-lane loop long for large ones
    double avg = c->nl.nnz_full > 0 ? (double)c->nl.nnz_full / c->n : 32.0;
    long waves1 = c->n / 64;
    int lpp = 1;
    while (lpp < 16 && waves1 * lpp < 4096 && avg / lpp >= 8) lpp *= 2;
    if (lpp == 2) lpp = 4;
    (void)kind;
    return lpp;
}

static void ensure_pp(sx_ctx *c)
{
    if (c->ecp_pp_cap < c->n) {
        if (c->d_ecp_pp) HIPCHK(hipFree(c->d_ecp_pp));
        if (c->d_ecp_scr) HIPCHK(hipFree(c->d_ecp_scr));
        HIPCHK(hipMalloc(&c->d_ecp_pp, sizeof(double) * PP_STRIDE * (size_t)c->n));
        HIPCHK(hipMalloc(&c->d_ecp_scr, sizeof(float) * ECP_SCR_STRIDE * (size_t)c->n));
        c->ecp_pp_cap = c->n;
        c->have_graph = false;
    }
}

static void ensure_pair_buffers(sx_ctx *c)
{
    long need = c->nl.cap_half;
    if (need > c->pair_cap) {
        if (c->d_pairbuf) HIPCHK(hipFree(c->d_pairbuf));
        if (c->d_live) HIPCHK(hipFree(c->d_live));
        if (c->d_work) HIPCHK(hipFree(c->d_work));
        HIPCHK(hipMalloc(&c->d_pairbuf, sizeof(double) * 10 * need));
        HIPCHK(hipMalloc(&c->d_live, need));
        HIPCHK(hipMalloc(&c->d_work, sizeof(int) * (need + 64)));
        c->pair_cap = need;
        c->have_graph = false;
    }
    if (c->ld > c->ecp_slot_n) {
        if (c->d_ecp_slot) HIPCHK(hipFree(c->d_ecp_slot));
        HIPCHK(hipMalloc(&c->d_ecp_slot, sizeof(int2) * ECP_SLOT_LINE * c->ld));
        HIPCHK(hipMemsetAsync(c->d_ecp_slot, 0, sizeof(int2) * ECP_SLOT_LINE * c->ld, c->stream));
        c->ecp_slot_n = c->ld;
        c->have_graph = false;
    }
    ensure_pp(c);
}

// learned warm start: predictions for the work list of the current call
const float *sx_ecp_mlp_predict(sx_ctx *c);
// learned solve stage (network and solve in one kernel)
void sx_ecp_solve_learned(sx_ctx *c, int variant, const EcpPar &par, long *stats);

static EcpScreen make_screen(sx_ctx *c, int kind, const Params &P)
{
    EcpScreen sc;
    sc.fvar = kind == SX_ECP ? 0 : kind == SX_ECPR ? 1 : kind == SX_ECPW ? 2 : kind == SX_ECPC ? 3 : 4;
    sc.par.cutoff = sc.fvar == 1 ? 0.0 : P.p[0];
    sc.par.wf = (sc.fvar == 2 || sc.fvar == 4) ? P.p[1] : 0.0;
    sc.par.chir = sc.fvar == 3 ? P.p[1] : (sc.fvar == 4 ? P.p[2] : 0.0);
    sc.variant = sc.fvar == 1 ? 1 : 0;
    sc.cutoff = sc.par.cutoff + sc.par.wf;
    sc.periodic = c->periodic;
    for (int d = 0; d < 3; ++d) {
        sc.L[d] = c->box[d];
        sc.invL[d] = 1.0 / c->box[d];
        sc.Lf[d] = (float)sc.L[d];
        sc.invLf[d] = (float)sc.invL[d];
    }
    return sc;
}

struct EcpLaunch {
    long cap;
    bool stats;   // count solved pairs / iterations (profiling)
};

static EcpLaunch ecp_launch_shape(sx_ctx *c, bool stats)
{
    EcpLaunch L;
    L.cap = c->nl.cap_half;
    L.stats = stats;
    return L;
}

static void ecp_stage(sx_ctx *c, int stage, int kind, const EcpScreen &sc, const EcpLaunch &L)
{
    NList &nl = c->nl;
    DState s = c->dstate();
    int n = c->n;
    long ld = 0;
    const int *nnz_half = c->d_flag + 2;
    switch (stage) {
    case 0:
        hipLaunchKernelGGL(k_ecp_pp, dim3(sx_blocks(n, 256)), dim3(256), 0, c->stream, s, c->d_ecp_pp, c->d_ecp_scr,
                           (int *)nullptr);
        break;
    case 1:
        hipLaunchKernelGGL(k_ecp_screen, dim3(ECP_NSEG), dim3(256), 0, c->stream, (const float *)c->d_ecp_scr, nl.half_i, nl.half_j, nnz_half, sc, c->d_live, c->d_work, c->d_segcount);
        break;
    case 2: {
        long *stats = L.stats ? c->d_counts : nullptr;
        if (c->ecp_solver == SX_ECP_LEARNED) {
            sx_ecp_solve_learned(c, sc.fvar, sc.par, stats);
            break;
        }
        const float *lam0 = nullptr;
#define SX_ECP_SOLVE(V)                                                                                          \
    hipLaunchKernelGGL(k_ecp_solve<V>, dim3(ECP_NSEG), dim3(ECP_SOLVE_WG), 0, c->stream, s, c->d_ecp_pp, ld, \
                       nl.half_i, nl.half_j, c->d_work, c->d_segcount, nnz_half, sc, c->ecp_solver, lam0,         \
                       c->d_pairbuf, c->d_live, stats, EcpSlots{c->d_ecp_slot})
        switch (sc.fvar) {
        case 0: SX_ECP_SOLVE(0); break;
        case 1: SX_ECP_SOLVE(1); break;
        case 2: SX_ECP_SOLVE(2); break;
        case 3: SX_ECP_SOLVE(3); break;
        default: SX_ECP_SOLVE(4); break;
        }
#undef SX_ECP_SOLVE
        break;
    }
    case 3:
        hipLaunchKernelGGL(k_ecp_apply, dim3(sx_blocks(n, 256)), dim3(256), 0, c->stream, s, sx_ecp_apply_args(c),
                           (const unsigned char *)c->d_live, c->d_pe_acc);
        break;
    }
}

// contact parameter only, for the solver comparison: MODE 0 the reference's
// Brent, 1 Newton from psi, 2 the learned solve (warm start lam0; pairs whose
// warm start proves them out of range end there), 3 the reference's Brent
// with its tolerances tightened to atol = rtol = 1e-10.  iters (may be NULL):
// [0] += iterations, [1] += the wave's maximum iterations of each round of
// 64 pairs, [2] += rounds, [3] += pairs proven out of range
template <int MODE>
__global__ void __launch_bounds__(ECP_SOLVE_THREADS) k_ecp_lambda(DState s, const double *pp, const int *half_i,
                                                                  const int *half_j, const int *work,
                                                                  const int *segcount, const int *nnz_half,
                                                                  EcpScreen sc, const float *lam0, double *lam_out,
                                                                  long *iters)
{
    long start, end;
    ecp_segment(*nnz_half, blockIdx.x, start, end);
    const int cnt = segcount[blockIdx.x];
    long my_it = 0, my_wmax = 0, rounds = 0, my_skip = 0;
    for (int base = 0; base < cnt; base += ECP_SOLVE_THREADS, ++rounds) {
        int u = base + threadIdx.x;
        int it = 0;
        if (u < cnt) {
            long t = start + u;
            int k = work[t];
            int a = half_i[k], b = half_j[k];
            double RIJ[3];
            pair_rij(s, a, b, RIJ);
            Sym3 A = load_sym(pp, 0, PP_AR, a), B = load_sym(pp, 0, PP_AR, b);
            double l;
            if (MODE == 0) {
                l = ecp_brent(A, B, RIJ, ecp_optim_dist(ECP_PSI, A, B, RIJ), it);
            } else if (MODE == 3) {
                l = ecp_brent(A, B, RIJ, ecp_optim_dist(ECP_PSI, A, B, RIJ), it, 1e-10, 1e-10);
            } else if (MODE == 1) {
                int ok = 0;
                l = ecp_lambda_newton(A, B, RIJ, ECP_PSI, 1e-10, it, ok);
            } else {
                // F and its derivatives at the warm start serve both the
                // out-of-range proof and the first Newton step
                l = fmin(fmax((double)lam0[t], 1e-6), 1.0 - 1e-6);
                double rsq = RIJ[0] * RIJ[0] + RIJ[1] * RIJ[1] + RIJ[2] * RIJ[2];
                double d[3];
                ecp_F_derivs(l, A, B, RIJ, d[0], d[1], d[2]);
                if (ecp_zero_proven(d[0], rsq, sc.variant, sc.cutoff)) {
                    ++my_skip;
                    l = -1.0;   // out of range: no contact parameter needed
                } else {
                    int ok = 0;
                    l = ecp_lambda_newton(A, B, RIJ, l, 1e-10, it, ok, d);
                }
            }
            lam_out[t] = l;
        }
        my_it += it;
        int wm = it;
        for (int off = SX_WAVE / 2; off > 0; off >>= 1) wm = max(wm, __shfl_xor(wm, off, SX_WAVE));
        my_wmax += wm;
    }
    (void)end;
    if (!iters) return;
    for (int off = SX_WAVE / 2; off > 0; off >>= 1) {
        my_it += __shfl_xor(my_it, off, SX_WAVE);
        my_skip += __shfl_xor(my_skip, off, SX_WAVE);
    }
    if ((threadIdx.x & (SX_WAVE - 1)) == 0 && cnt) {
        atomicAdd((unsigned long long *)&iters[0], (unsigned long long)my_it);
        atomicAdd((unsigned long long *)&iters[1], (unsigned long long)my_wmax);
        atomicAdd((unsigned long long *)&iters[2], (unsigned long long)rounds);
        if (my_skip) atomicAdd((unsigned long long *)&iters[3], (unsigned long long)my_skip);
    }
}

// per-particle records and the screened work list of kind on the current
// state (the first two stages of a force call)
void sx_ecp_prepare(sx_ctx *c, int kind, const double *p, int np)
{
    if (!is_ecp_kind(kind)) throw sx_error{SX_ERR_ARG, "an ECP kind is required"};
    if (!c->nl.init) throw sx_error{SX_ERR_STATE, "neighbour list not initialised (sx_nlist_init)"};
    check_params(kind, np);
    sx_nlist_want_half(c);
    Params P = make_params(c, kind, p, np);
    ensure_pair_buffers(c);
    EcpScreen sc = make_screen(c, kind, P);
    EcpLaunch L = ecp_launch_shape(c, false);
    ecp_stage(c, 0, kind, sc, L);
    ecp_stage(c, 1, kind, sc, L);
}

// ---------------------------------------------------------------------------
// reaction field of the point dipoles (module dipole, force.f90:2669-2767)
// ---------------------------------------------------------------------------
// per particle i, over its list entries j > i within the cutoff (the pairs
// the reference's loop hands to dipole_calculate_forces with i first): DIP(i)
// = i's scaled axis when there is one; the field increments of those pairs
// are summed per particle, then over particles by the last workgroup
__global__ void __launch_bounds__(256) k_dipole_rf(DState s, const int *row_start, const int *nbr, double cutoff,
                                                   double dip_coeff, double *dip, double *part, unsigned *ticket,
                                                   double *rfl)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    if (i < s.n) {
        bool any = false;
        for (int k = row_start[i]; k < row_start[i + 1]; ++k) {
            int j = nbr[k];
            if (j <= i) continue;
            double r[3], rsq, rmag;
            if (!dip_in_range(s, i, j, cutoff, r, rsq, rmag)) continue;
            any = true;
            double u2[3];
            dip_axis(s, j, u2);
            for (int q = 0; q < 3; ++q) v[q] = v[q] + u2[q] * dip_coeff;
        }
        if (any) {
            double u1[3];
            dip_axis(s, i, u1);
            for (int q = 0; q < 3; ++q) dip[(long)q * s.n + i] = u1[q];
        }
    }
    SxRedIds ids;
    ids.id[0] = 0; ids.id[1] = 1; ids.id[2] = 2; ids.n = 3; ids.accumulate = 1;
    sx_reduce_last_block<3>(v, part, ticket, rfl, ids);
}

// eval_reaction_field (force.f90:2753-2766) per particle: the energy term
// goes to the particle's energy slot
__global__ void k_eval_rf(DState s, const double *dip, const double *rfl, double *pe_acc)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n || s.f[SX_CHARGE][i] == 0.0) return;
    long n = s.n;
    double dx = dip[i], dy = dip[n + i], dz = dip[2 * n + i];
    double rx = rfl[0], ry = rfl[1], rz = rfl[2];
    pe_acc[i] = pe_acc[i] + 0.5 * (dx * rx + dy * ry + dz * rz);
    s.f[SX_TX][i] = s.f[SX_TX][i] - (dy * rz - dz * ry);
    s.f[SX_TY][i] = s.f[SX_TY][i] - (dz * rx - dx * rz);
    s.f[SX_TZ][i] = s.f[SX_TZ][i] - (dx * ry - dy * rx);
}

static void need_rf(sx_ctx *c)
{
    if (!c->d_dip || c->dip_cap != c->n)
        throw sx_error{SX_ERR_STATE, "reaction field not initialised (sx_init_reaction_field)"};
}

static void dipole_side_effects(sx_ctx *c, const Params &P)
{
    need_rf(c);
    double cutoff = P.p[0], dielectric = P.p[1];
    double dip_coeff = (2.0 * (dielectric - 1.0)) / (2.0 * dielectric + 1.0);
    dip_coeff = dip_coeff / (cutoff * cutoff * cutoff);
    int nb = sx_blocks(c->n, 256);
    sx_ensure_part(c, nb, 3);
    hipLaunchKernelGGL(k_dipole_rf, dim3(nb), dim3(256), 0, c->stream, c->dstate(), c->nl.row_start, c->nl.nbr,
                       cutoff, dip_coeff, c->d_dip, c->d_part, c->d_ticket, c->d_rfl);
}

void sx_eval_reaction_field_impl(sx_ctx *c)
{
    need_rf(c);
    hipLaunchKernelGGL(k_eval_rf, dim3(sx_blocks(c->n, 256)), dim3(256), 0, c->stream, c->dstate(), c->d_dip,
                       c->d_rfl, c->d_pe_acc);
    if (!c->defer_pe) sx_pe_total_impl(c);
}

static void ecp_forces_impl(sx_ctx *c, int kind, const Params &P)
{
    ensure_pair_buffers(c);
    EcpScreen sc = make_screen(c, kind, P);
    EcpLaunch L = ecp_launch_shape(c, false);
    // the whole-step driver writes the records in its initial-step kernel,
    // and may leave the apply stage to its final-step kernel
    for (int stage = c->ecp_pp_by_step ? 1 : 0; stage < (c->ecp_apply_in_final ? 3 : 4); ++stage)
        ecp_stage(c, stage, kind, sc, L);
}

void sx_ecp_records(sx_ctx *c, double **pp, float **scr)
{
    ensure_pp(c);
    *pp = c->d_ecp_pp;
    *scr = c->d_ecp_scr;
}

template <int K1, int K2, typename T>
static void launch_gather2_t(sx_ctx *c, const Params &P1, const Params &P2, int lpp)
{
    const int nb = sx_blocks((long)c->n * lpp, 256);
    DState s = c->dstate();
    switch (lpp) {
    case 1: hipLaunchKernelGGL((k_gather2<K1, K2, 1, T>), dim3(nb), dim3(256), 0, c->stream, s, c->nl.row_start, c->nl.nbr, P1, P2, c->d_pe_acc); break;
    case 4: hipLaunchKernelGGL((k_gather2<K1, K2, 4, T>), dim3(nb), dim3(256), 0, c->stream, s, c->nl.row_start, c->nl.nbr, P1, P2, c->d_pe_acc); break;
    case 8: hipLaunchKernelGGL((k_gather2<K1, K2, 8, T>), dim3(nb), dim3(256), 0, c->stream, s, c->nl.row_start, c->nl.nbr, P1, P2, c->d_pe_acc); break;
    default: hipLaunchKernelGGL((k_gather2<K1, K2, 16, T>), dim3(nb), dim3(256), 0, c->stream, s, c->nl.row_start, c->nl.nbr, P1, P2, c->d_pe_acc); break;
    }
}

template <int K1, int K2>
static void launch_gather2(sx_ctx *c, const Params &P1, const Params &P2, int lpp)
{
    if (c->pair_f32) launch_gather2_t<K1, K2, float>(c, P1, P2, lpp);
    else launch_gather2_t<K1, K2, double>(c, P1, P2, lpp);
}

// two pair loops of the step driver in one pass where both kinds are WCA and
// Debye-Hueckel (the charged-chain step); false: not such a pair
bool sx_forces_two_impl(sx_ctx *c, int k1, const double *p1, int np1, int k2, const double *p2, int np2)
{
    const bool wd = k1 == SX_WCA && k2 == SX_DH, dw = k1 == SX_DH && k2 == SX_WCA;
    if (!wd && !dw) return false;
    check_params(k1, np1);
    check_params(k2, np2);
    if (!c->nl.init) throw sx_error{SX_ERR_STATE, "neighbour list not initialised (sx_nlist_init)"};
    const Params P1 = make_params(c, k1, p1, np1), P2 = make_params(c, k2, p2, np2);
    const int lpp = choose_lpp(c, k1);   // the same for every kind
    if (wd) launch_gather2<SX_WCA, SX_DH>(c, P1, P2, lpp);
    else launch_gather2<SX_DH, SX_WCA>(c, P1, P2, lpp);
    if (!c->defer_pe) sx_pe_total_impl(c);
    return true;
}

void sx_forces_impl(sx_ctx *c, int kind, const double *p, int np)
{
    check_params(kind, np);
    if (!c->nl.init) throw sx_error{SX_ERR_STATE, "neighbour list not initialised (sx_nlist_init)"};
    Params P = make_params(c, kind, p, np);
    if (is_ecp_kind(kind)) {
        ecp_forces_impl(c, kind, P);
    } else {
        int lpp = choose_lpp(c, kind);
        switch (kind) {
        case SX_LJ: launch_gather<SX_LJ, false>(c, P, lpp); break;
        case SX_LJ12: launch_gather<SX_LJ12, false>(c, P, lpp); break;
        case SX_LJW: launch_gather<SX_LJW, false>(c, P, lpp); break;
        case SX_WCA: launch_gather<SX_WCA, false>(c, P, lpp); break;
        case SX_MORSE: launch_gather<SX_MORSE, false>(c, P, lpp); break;
        case SX_DH: launch_gather<SX_DH, false>(c, P, lpp); break;
        case SX_COULOMB: launch_gather<SX_COULOMB, false>(c, P, lpp); break;
        case SX_GB: launch_gather_gb<0>(c, P, lpp); break;
        case SX_GBR: launch_gather_gb<1>(c, P, lpp); break;
        case SX_GBW: launch_gather_gb<2>(c, P, lpp); break;
        case SX_GBC: launch_gather_gb<3>(c, P, lpp); break;
        case SX_GBWC: launch_gather_gb<4>(c, P, lpp); break;
        case SX_DIPOLE:
            launch_gather<SX_DIPOLE, true>(c, P, lpp);
            dipole_side_effects(c, P);
            break;
        }
    }
    // the step driver folds this reduction into the final-step kernel
    if (!c->defer_pe) sx_pe_total_impl(c);
}

extern "C" {

int sx_init_reaction_field(sx_ctx *c)
{
    return guard(c, [&] {
        need_system(c);
        if (c->dip_cap != c->n) {
            if (c->d_dip) HIPCHK(hipFree(c->d_dip));
            HIPCHK(hipMalloc(&c->d_dip, sizeof(double) * 3 * (size_t)c->n));
            if (!c->d_rfl) HIPCHK(hipMalloc(&c->d_rfl, sizeof(double) * 3));
            c->dip_cap = c->n;
            c->have_graph = false;
        }
        HIPCHK(hipMemsetAsync(c->d_dip, 0, sizeof(double) * 3 * (size_t)c->n, c->stream));
        HIPCHK(hipMemsetAsync(c->d_rfl, 0, sizeof(double) * 3, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int sx_eval_reaction_field(sx_ctx *c)
{
    return guard(c, [&] { need_system(c); sx_eval_reaction_field_impl(c); });
}

int sx_get_reaction_field(sx_ctx *c, double *dip, double rfl[3])
{
    return guard(c, [&] {
        need_system(c);
        need_rf(c);
        if (dip)
            HIPCHK(hipMemcpyAsync(dip, c->d_dip, sizeof(double) * 3 * (size_t)c->n, hipMemcpyDeviceToHost, c->stream));
        if (rfl) HIPCHK(hipMemcpyAsync(rfl, c->d_rfl, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int sx_forces(sx_ctx *c, int kind, const double *params, int nparams)
{
    return guard(c, [&] {
        need_system(c);
        if (is_ecp_kind(kind) && c->nl.init) sx_nlist_want_half(c);
        sx_forces_impl(c, kind, params, nparams);
    });
}

int sx_profile_ecp(sx_ctx *c, int kind, const double *params, int nparams, int nrep, double ms[4], double stats[3])
{
    return guard(c, [&] {
        need_system(c);
        if (!is_ecp_kind(kind)) throw sx_error{SX_ERR_ARG, "profile_ecp needs an ECP kind"};
        if (!c->nl.init) throw sx_error{SX_ERR_STATE, "neighbour list not initialised (sx_nlist_init)"};
        check_params(kind, nparams);
        sx_nlist_want_half(c);
        Params P = make_params(c, kind, params, nparams);
        ensure_pair_buffers(c);
        EcpScreen sc = make_screen(c, kind, P);
        // timed with the launch shape of a normal step (no counters), then one
        // extra pass with the solve counters on for the pair statistics
        EcpLaunch L = ecp_launch_shape(c, false);
        EcpLaunch Ls = ecp_launch_shape(c, true);
        // as inside a replayed step: a stage launched nrep times in a
        // captured graph, replayed once untimed and three times timed, the
        // time per launch the median replay over nrep (the graph's node
        // dispatch included, about 1 us).  The apply stage consumes the slots
        // the solve takes, so it is timed as screen + solve + apply less the
        // screen and the solve alone.
        // the apply stage adds into the forces, torques and pair energies:
        // they are kept aside here and put back after the timed passes
        const size_t ld = (size_t)c->ld;
        double *keep = nullptr;
        HIPCHK(hipMalloc(&keep, sizeof(double) * 7 * ld));
        double *const fld = c->d_fields;
        auto hold = [&](bool back) {
            double *spans[3] = {fld + (size_t)SX_FX * ld, fld + (size_t)SX_TX * ld, c->d_pe_acc};
            const size_t len[3] = {3 * ld, 3 * ld, ld};
            size_t off = 0;
            for (int k = 0; k < 3; ++k) {
                if (back) HIPCHK(hipMemcpyAsync(spans[k], keep + off, sizeof(double) * len[k], hipMemcpyDeviceToDevice, c->stream));
                else HIPCHK(hipMemcpyAsync(keep + off, spans[k], sizeof(double) * len[k], hipMemcpyDeviceToDevice, c->stream));
                off += len[k];
            }
        };
        hold(false);
        for (int stage = 0; stage < 4; ++stage) ecp_stage(c, stage, kind, sc, L);   // buffers sized
        HIPCHK(hipStreamSynchronize(c->stream));
        hipEvent_t ev[2];
        for (auto &e : ev) HIPCHK(hipEventCreate(&e));
        std::vector<double> times[4];
        long dev0[8], dev1[8];
        for (int q = 0; q < 4; ++q) {
            hipGraph_t g = nullptr;
            hipGraphExec_t ge = nullptr;
            HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            for (int r = 0; r < nrep; ++r) {
                if (q < 3) ecp_stage(c, q, kind, sc, L);
                else
                    for (int st = 1; st < 4; ++st) ecp_stage(c, st, kind, sc, L);
            }
            HIPCHK(hipStreamEndCapture(c->stream, &g));
            HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            for (int r = -1; r < 3; ++r) {
                HIPCHK(hipEventRecord(ev[0], c->stream));
                HIPCHK(hipGraphLaunch(ge, c->stream));
                HIPCHK(hipEventRecord(ev[1], c->stream));
                HIPCHK(hipEventSynchronize(ev[1]));
                float t = 0;
                HIPCHK(hipEventElapsedTime(&t, ev[0], ev[1]));
                if (r >= 0) times[q].push_back(t / nrep);
            }
            HIPCHK(hipGraphExecDestroy(ge));
            HIPCHK(hipGraphDestroy(g));
            // the solves alone left their slot counts standing
            if (q == 2) ecp_clear_slots(c);
        }
        // q = 3 timed screen + solve + apply; apply = that less the two alone
        for (int r = 0; r < 3; ++r) times[3][r] = std::max(0.0, times[3][r] - times[1][r] - times[2][r]);
        HIPCHK(hipMemcpyAsync(dev0, c->d_counts, sizeof(long) * 8, hipMemcpyDeviceToHost, c->stream));
        for (int stage = 0; stage < 4; ++stage) ecp_stage(c, stage, kind, sc, Ls);
        HIPCHK(hipMemcpyAsync(dev1, c->d_counts, sizeof(long) * 8, hipMemcpyDeviceToHost, c->stream));
        int sizes[8];
        HIPCHK(hipMemcpyAsync(sizes, c->d_flag, sizeof(int) * 8, hipMemcpyDeviceToHost, c->stream));
        hold(true);
        HIPCHK(hipStreamSynchronize(c->stream));
        hipFree(keep);
        for (auto &e : ev) hipEventDestroy(e);
        for (int q = 0; q < 4; ++q) {
            std::vector<double> &v = times[q];
            if (v.empty()) { ms[q] = 0.0; continue; }
            std::sort(v.begin(), v.end());
            ms[q] = v.size() % 2 ? v[v.size() / 2] : 0.5 * (v[v.size() / 2 - 1] + v[v.size() / 2]);
        }
        stats[0] = sizes[2];
        stats[1] = (double)(dev1[0] - dev0[0]);
        stats[2] = (double)(dev1[1] - dev0[1]);
        sx_check_device_error(c);
    });
}

int sx_set_ecp_solver(sx_ctx *c, int solver)
{
    return guard(c, [&] {
        if (solver < SX_ECP_BRENT_REFERENCE || solver > SX_ECP_LEARNED) throw sx_error{SX_ERR_ARG, "bad solver"};
        if (solver == SX_ECP_LEARNED && !c->mlp) throw sx_error{SX_ERR_STATE, "no lambda* network loaded"};
        c->ecp_solver = solver;
        c->have_graph = false;
    });
}

int sx_set_pair_precision(sx_ctx *c, int bits)
{
    return guard(c, [&] {
        if (bits != 32 && bits != 64) throw sx_error{SX_ERR_ARG, "pair precision must be 32 or 64 bits"};
        if (c->pair_f32 != (bits == 32)) c->have_graph = false;
        c->pair_f32 = bits == 32;
    });
}

int sx_get_pair_precision(sx_ctx *c, int *bits)
{
    return guard(c, [&] { *bits = c->pair_f32 ? 32 : 64; });
}

int sx_ecp_stats(sx_ctx *c, long out[4])
{
    return guard(c, [&] {
        long dev[8];
        HIPCHK(hipMemcpyAsync(dev, c->d_counts, sizeof(long) * 8, hipMemcpyDeviceToHost, c->stream));
        int sizes[8];
        HIPCHK(hipMemcpyAsync(sizes, c->d_flag, sizeof(int) * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        out[0] = sizes[2];     // half pairs screened (last call)
        out[1] = dev[0];       // pairs solved (cumulative)
        out[2] = dev[1];       // solver iterations (cumulative)
        out[3] = dev[2];       // fallbacks (cumulative)
    });
}

int sx_pair_eval(sx_ctx *c, int kind, const double *params, int nparams, long np, const int *i1, const int *j1,
                 double *pe, double *f, double *ti, double *tj, double *lambda)
{
    return guard(c, [&] {
        need_system(c);
        check_params(kind, nparams);
        if (np <= 0) return;
        for (long t = 0; t < np; ++t)
            if (i1[t] < 1 || i1[t] > c->n || j1[t] < 1 || j1[t] > c->n)
                throw sx_error{SX_ERR_ARG, "Error: Atom index does not exist in system"};
        Params P = make_params(c, kind, params, nparams);
        std::vector<int> hi(np), hj(np);
        for (long t = 0; t < np; ++t) { hi[t] = i1[t] - 1; hj[t] = j1[t] - 1; }
        int *di = nullptr, *dj = nullptr;
        double *dout = nullptr;
        HIPCHK(hipMalloc(&di, sizeof(int) * np));
        HIPCHK(hipMalloc(&dj, sizeof(int) * np));
        HIPCHK(hipMalloc(&dout, sizeof(double) * 11 * np));
        HIPCHK(hipMemcpyAsync(di, hi.data(), sizeof(int) * np, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(dj, hj.data(), sizeof(int) * np, hipMemcpyHostToDevice, c->stream));
        DState s = c->dstate();
        if (is_ecp_kind(kind)) {
            ensure_pp(c);
            hipLaunchKernelGGL(k_ecp_pp, dim3(sx_blocks(c->n, 256)), dim3(256), 0, c->stream, s, c->d_ecp_pp,
                               c->d_ecp_scr, (int *)nullptr);
        }
        int solver = c->ecp_solver == SX_ECP_LEARNED ? SX_ECP_CONVERGED : c->ecp_solver;
        long ld = 0;
        dim3 g(sx_blocks(np, 128)), b(128);
        switch (kind) {
        case SX_LJ: hipLaunchKernelGGL(k_pair_eval<SX_LJ>, g, b, 0, c->stream, s, np, di, dj, P, c->d_ecp_pp, ld, solver, dout); break;
        case SX_LJ12: hipLaunchKernelGGL(k_pair_eval<SX_LJ12>, g, b, 0, c->stream, s, np, di, dj, P, c->d_ecp_pp, ld, solver, dout); break;
        case SX_LJW: hipLaunchKernelGGL(k_pair_eval<SX_LJW>, g, b, 0, c->stream, s, np, di, dj, P, c->d_ecp_pp, ld, solver, dout); break;
        case SX_WCA: hipLaunchKernelGGL(k_pair_eval<SX_WCA>, g, b, 0, c->stream, s, np, di, dj, P, c->d_ecp_pp, ld, solver, dout); break;
        case SX_MORSE: hipLaunchKernelGGL(k_pair_eval<SX_MORSE>, g, b, 0, c->stream, s, np, di, dj, P, c->d_ecp_pp, ld, solver, dout); break;
        case SX_DH: hipLaunchKernelGGL(k_pair_eval<SX_DH>, g, b, 0, c->stream, s, np, di, dj, P, c->d_ecp_pp, ld, solver, dout); break;
        case SX_GB: hipLaunchKernelGGL(k_pair_eval<SX_GB>, g, b, 0, c->stream, s, np, di, dj, P, c->d_ecp_pp, ld, solver, dout); break;
        case SX_GBR: hipLaunchKernelGGL(k_pair_eval<SX_GBR>, g, b, 0, c->stream, s, np, di, dj, P, c->d_ecp_pp, ld, solver, dout); break;
        case SX_ECP: hipLaunchKernelGGL(k_pair_eval<SX_ECP>, g, b, 0, c->stream, s, np, di, dj, P, c->d_ecp_pp, ld, solver, dout); break;
        case SX_ECPR: hipLaunchKernelGGL(k_pair_eval<SX_ECPR>, g, b, 0, c->stream, s, np, di, dj, P, c->d_ecp_pp, ld, solver, dout); break;
        case SX_GBW: hipLaunchKernelGGL(k_pair_eval<SX_GBW>, g, b, 0, c->stream, s, np, di, dj, P, c->d_ecp_pp, ld, solver, dout); break;
        case SX_GBC: hipLaunchKernelGGL(k_pair_eval<SX_GBC>, g, b, 0, c->stream, s, np, di, dj, P, c->d_ecp_pp, ld, solver, dout); break;
        case SX_GBWC: hipLaunchKernelGGL(k_pair_eval<SX_GBWC>, g, b, 0, c->stream, s, np, di, dj, P, c->d_ecp_pp, ld, solver, dout); break;
        case SX_ECPW: hipLaunchKernelGGL(k_pair_eval<SX_ECPW>, g, b, 0, c->stream, s, np, di, dj, P, c->d_ecp_pp, ld, solver, dout); break;
        case SX_ECPC: hipLaunchKernelGGL(k_pair_eval<SX_ECPC>, g, b, 0, c->stream, s, np, di, dj, P, c->d_ecp_pp, ld, solver, dout); break;
        case SX_ECPWC: hipLaunchKernelGGL(k_pair_eval<SX_ECPWC>, g, b, 0, c->stream, s, np, di, dj, P, c->d_ecp_pp, ld, solver, dout); break;
        case SX_COULOMB: hipLaunchKernelGGL(k_pair_eval<SX_COULOMB>, g, b, 0, c->stream, s, np, di, dj, P, c->d_ecp_pp, ld, solver, dout); break;
        case SX_DIPOLE: hipLaunchKernelGGL(k_pair_eval<SX_DIPOLE>, g, b, 0, c->stream, s, np, di, dj, P, c->d_ecp_pp, ld, solver, dout); break;
        }
        std::vector<double> h(11 * np);
        HIPCHK(hipMemcpyAsync(h.data(), dout, sizeof(double) * 11 * np, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        hipFree(di); hipFree(dj); hipFree(dout);
        for (long t = 0; t < np; ++t) {
            const double *r = &h[11 * t];
            if (pe) pe[t] = r[0];
            for (int k = 0; k < 3; ++k) {
                if (f) f[3 * t + k] = r[1 + k];
                if (ti) ti[3 * t + k] = r[4 + k];
                if (tj) tj[3 * t + k] = r[7 + k];
            }
            if (lambda) lambda[t] = r[10];
        }
    });
}

}  // extern "C"

// Contact-parameter solvers on the current state's ECP work list, each timed
// alone with HIP events over nrep runs: the reference's Brent, Brent with
// tolerances tightened to 1e-10, Newton from psi (plain) and Newton from the
// network's warm start (learned), plus the full solve stage (lambda +
// force/torque) of the converged and learned modes.  out[19]:
//  [0] live pairs  [1] ms Brent  [2] ms plain Newton  [3] ms network
//  [4] ms warm Newton  [5] ms full plain solve  [6] ms full learned solve
//  (network included)  [7] ms tight Brent  [8] iterations/pair plain
//  [9] iterations/pair warm  [10] mean per-wave maximum iterations plain
//  [11] same, warm  [12] mean |lambda0 - lambda*|  [13] max  [14] rms
//  [15] max |lambda_warm - lambda_plain|  [16] iterations/pair tight Brent
//  [17] mean per-wave maximum iterations, tight Brent  [18] pairs the
//  learned solve proved out of range from its warm start alone
extern "C" int sx_ecp_compare_solvers(sx_ctx *c, int kind, const double *params, int nparams, int nrep,
                                      double out[19])
{
    return guard(c, [&] {
        need_system(c);
        if (!c->mlp) throw sx_error{SX_ERR_STATE, "lambda* network not initialised (sx_mlp_init)"};
        if (nrep < 1) nrep = 1;
        sx_ecp_prepare(c, kind, params, nparams);
        Params P = make_params(c, kind, params, nparams);
        EcpScreen sc = make_screen(c, kind, P);
        NList &nl = c->nl;
        DState s = c->dstate();
        const int *nnz_half = c->d_flag + 2;
        long cap = nl.cap_half;
        double *lam_b = nullptr, *lam_p = nullptr, *lam_w = nullptr;
        long *cnt = nullptr;
        HIPCHK(hipMalloc(&lam_b, sizeof(double) * cap));
        HIPCHK(hipMalloc(&lam_p, sizeof(double) * cap));
        HIPCHK(hipMalloc(&lam_w, sizeof(double) * cap));
        HIPCHK(hipMalloc(&cnt, sizeof(long) * 12));
        HIPCHK(hipMemsetAsync(cnt, 0, sizeof(long) * 12, c->stream));
        const float *lam0 = sx_ecp_mlp_predict(c);
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        auto timed = [&](auto &&launch) {
            launch();   // warm
            HIPCHK(hipEventRecord(e0, c->stream));
            for (int r = 0; r < nrep; ++r) launch();
            HIPCHK(hipEventRecord(e1, c->stream));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            return (double)ms / nrep;
        };
        dim3 g(ECP_NSEG), b(ECP_SOLVE_THREADS);
        double t_brent = timed([&] {
            hipLaunchKernelGGL(k_ecp_lambda<0>, g, b, 0, c->stream, s, c->d_ecp_pp, nl.half_i, nl.half_j, c->d_work,
                               c->d_segcount, nnz_half, sc, lam0, lam_b, (long *)nullptr);
        });
        double t_brent_tight = timed([&] {
            hipLaunchKernelGGL(k_ecp_lambda<3>, g, b, 0, c->stream, s, c->d_ecp_pp, nl.half_i, nl.half_j, c->d_work,
                               c->d_segcount, nnz_half, sc, lam0, lam_b, (long *)nullptr);
        });
        double t_plain = timed([&] {
            hipLaunchKernelGGL(k_ecp_lambda<1>, g, b, 0, c->stream, s, c->d_ecp_pp, nl.half_i, nl.half_j, c->d_work,
                               c->d_segcount, nnz_half, sc, lam0, lam_p, (long *)nullptr);
        });
        double t_mlp = timed([&] { sx_ecp_mlp_predict(c); });
        double t_warm = timed([&] {
            hipLaunchKernelGGL(k_ecp_lambda<2>, g, b, 0, c->stream, s, c->d_ecp_pp, nl.half_i, nl.half_j, c->d_work,
                               c->d_segcount, nnz_half, sc, lam0, lam_w, (long *)nullptr);
        });
        auto full = [&](int solver, const float *l0) {
#define SX_ECP_FULL(V)                                                                                            \
    hipLaunchKernelGGL(k_ecp_solve<V>, g, dim3(ECP_SOLVE_WG), 0, c->stream, s, c->d_ecp_pp, 0L, nl.half_i,      \
                       nl.half_j, c->d_work,                                                                     \
                       c->d_segcount, nnz_half, sc, solver, l0, c->d_pairbuf, c->d_live, (long *)nullptr,    \
                       EcpSlots{c->d_ecp_slot})
            switch (sc.fvar) {
            case 0: SX_ECP_FULL(0); break;
            case 1: SX_ECP_FULL(1); break;
            case 2: SX_ECP_FULL(2); break;
            case 3: SX_ECP_FULL(3); break;
            default: SX_ECP_FULL(4); break;
            }
#undef SX_ECP_FULL
        };
        double t_full_plain = timed([&] { full(SX_ECP_CONVERGED, nullptr); });
        double t_full_learned = timed([&] { sx_ecp_solve_learned(c, sc.fvar, sc.par, nullptr); });
        // iteration counts (4 counters per solver)
        hipLaunchKernelGGL(k_ecp_lambda<3>, g, b, 0, c->stream, s, c->d_ecp_pp, nl.half_i, nl.half_j, c->d_work,
                           c->d_segcount, nnz_half, sc, lam0, lam_b, cnt + 0);
        hipLaunchKernelGGL(k_ecp_lambda<1>, g, b, 0, c->stream, s, c->d_ecp_pp, nl.half_i, nl.half_j, c->d_work,
                           c->d_segcount, nnz_half, sc, lam0, lam_p, cnt + 4);
        hipLaunchKernelGGL(k_ecp_lambda<2>, g, b, 0, c->stream, s, c->d_ecp_pp, nl.half_i, nl.half_j, c->d_work,
                           c->d_segcount, nnz_half, sc, lam0, lam_w, cnt + 8);
        long hc[12];
        int nnz = 0;
        std::vector<int> seg(ECP_NSEG);
        std::vector<double> hp(cap), hw(cap);
        std::vector<float> h0(cap);
        HIPCHK(hipMemcpyAsync(hc, cnt, sizeof(long) * 12, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(&nnz, nnz_half, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(seg.data(), c->d_segcount, sizeof(int) * ECP_NSEG, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(hp.data(), lam_p, sizeof(double) * cap, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(hw.data(), lam_w, sizeof(double) * cap, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(h0.data(), lam0, sizeof(float) * cap, hipMemcpyDeviceToHost, c->stream));
        // the solves above took apply-stage slots that no apply stage reads
        ecp_clear_slots(c);
        HIPCHK(hipStreamSynchronize(c->stream));
        long live = 0;
        double s_err = 0, s_err2 = 0, m_err = 0, m_dw = 0;
        long chunk = ((long)nnz + ECP_NSEG - 1) / ECP_NSEG;
        for (int q = 0; q < ECP_NSEG; ++q) {
            long st = std::min<long>((long)q * chunk, nnz);
            for (int u = 0; u < seg[q]; ++u) {
                long t = st + u;
                double e = std::fabs((double)h0[t] - hp[t]);
                s_err += e;
                s_err2 += e * e;
                m_err = std::max(m_err, e);
                if (hw[t] >= 0.0) m_dw = std::max(m_dw, std::fabs(hw[t] - hp[t]));   // pairs the learned solve solved
                ++live;
            }
        }
        double L = (double)std::max<long>(live, 1);
        auto per = [](long a, long b) { return b > 0 ? (double)a / (double)b : 0.0; };
        out[0] = (double)live;
        out[1] = t_brent; out[2] = t_plain; out[3] = t_mlp; out[4] = t_warm;
        out[5] = t_full_plain; out[6] = t_full_learned;
        out[7] = t_brent_tight;
        out[8] = hc[4] / L; out[9] = hc[8] / L;
        out[10] = per(hc[5], hc[6]); out[11] = per(hc[9], hc[10]);
        out[12] = s_err / L; out[13] = m_err; out[14] = std::sqrt(s_err2 / L); out[15] = m_dw;
        out[16] = hc[0] / L; out[17] = per(hc[1], hc[2]);
        out[18] = (double)hc[11];
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        HIPCHK(hipFree(lam_b));
        HIPCHK(hipFree(lam_p));
        HIPCHK(hipFree(lam_w));
        HIPCHK(hipFree(cnt));
        sx_check_device_error(c);
    });
}
