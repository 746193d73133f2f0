// Batched replicas: many independent small systems advanced together, each a
// contiguous range of the context's particle arrays with its own cubic
// periodic box, pair potential (LJ or Gay-Berne), parameters and Langevin
// temperature.  Per system and step the reference's loop (README.md:83-115):
// lgv_initial_step, initialize_forces, the pair loop over all pairs I < J of
// the system (the README's own double loop; the cutoff of each potential does
// the screening, so no neighbour list is built), lgv_final_step and
// calculate_thermo's KE / rotational KE / PE / T for that system.
//
// A step is three launches, captured once in a HIP graph and replayed:
//   k_batch_initial  one thread per particle: lgv_initial_step, then the
//                    particle's row of its system's table;
//   k_batch_forces   workgroups over (system, slice of its particles), largest
//                    work first, so a large system spreads over several CUs;
//                    each stages its system's table in LDS and gathers the
//                    forces, torques and pair energies of its slice;
//   k_batch_final    one workgroup per system: lgv_final_step and the thermo
//                    sums; the last workgroup advances the step counter.
// Forces are gathered per particle (LPP lanes per particle, combined with
// cross-lane shuffles in a fixed order, or one compacting wavefront per
// particle): no atomics, deterministic results.  The slice size and lanes per
// particle are picked by sx_batch_autotune from timed trial steps.
#include "sx_steps.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

struct BatchSys {
    int off, n, kind, temp_int;
    double L, T;
    double e1, e2, kk1, kk2;
    double p[4];
};

struct BatchState {
    int nsys = 0;
    std::vector<BatchSys> h_sys;
    BatchSys *d_sys = nullptr;
    int *d_sys_of = nullptr;       // system of each particle
    double *d_tab = nullptr;       // per-particle tables, TB_N doubles per particle
    double *d_pe = nullptr;        // per particle: its pair energies (pairs where it is I)
    int4 *d_blocks = nullptr;      // force workgroups: system, first, end particle
    int nblocks = 0, blocks_choice = -1, blocks_cap = 0;
    double *d_out = nullptr;       // per system: KE, KErot, PE, T
    unsigned *d_ticket = nullptr;
    int *d_step = nullptr;
    int choice = 6;
    int cap = 0;                   // largest system
    double dt = -1, damp = -1;     // constants currently uploaded
    int *d_order = nullptr;        // systems, most expensive first (one-launch step)
    unsigned *d_qctr = nullptr;    // its queue and done counters
    int ngrid = 0;                 // its workgroups: two per CU
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_choice = -1;
    double graph_dt = -1, graph_damp = -1;   // kernel arguments baked into the graph
    bool graph_f32 = false;                  // and the pair terms' precision
    int n_total = 0;                         // particle count and field storage at sx_batch_set
    const double *fields = nullptr;
};

// Three-launch shapes 0-7: (particles per force workgroup, lanes per
// particle); 64 lanes = one compacting wavefront per particle.  Shape 0 adds
// each particle's pair terms in the README loop's order.  Shapes 8 and 9:
// the whole step in one launch (k_batch_step) with workgroups of 256 and 512
// threads taking whole systems (P = 0 marks them).
static const int BATCH_SHAPES[][2] = {{64, 1},  {16, 4},  {32, 4}, {32, 8}, {64, 8},
                                      {16, 64}, {32, 64}, {64, 64}, {0, 256}, {0, 512}};
static const int BATCH_NSHAPES = 10;
static bool shape_fused(int choice) { return BATCH_SHAPES[choice][0] == 0; }

// Per-particle table of every system, written right after lgv_initial_step
// (one SoA block of TB_N rows per system, row length N_s): position, and for
// Gay-Berne the factors of gb_ that depend on one particle only (GbSelf) plus
// its reach, an upper bound on the distance at which the pair with that
// particle as I can interact:
// max(sig_x, sig_z) + cutoff - sigma_0 bounds gb_'s sigma + cutoff - sigma_0
// because sigma(u1, u2, r) never exceeds the larger of the two diameters.
enum { TB_X, TB_Y, TB_Z, TB_U0, TB_U1, TB_U2, TB_CHI, TB_XHI, TB_LPRP, TB_EPSZ, TB_REACH, TB_N };
static const int BATCH_LDS_BYTES = 64 * 1024;

// particle q of the table: position and (Gay-Berne) its own factors and reach
struct BatchSelf {
    double x, y, z, reach;
    GbSelf g;
};

// Gay-Berne factors of table particle q
__device__ __forceinline__ void batch_gb_self(const double *tab, int ld, int q, GbSelf &g)
{
    g.u[0] = tab[TB_U0 * ld + q]; g.u[1] = tab[TB_U1 * ld + q]; g.u[2] = tab[TB_U2 * ld + q];
    g.chi = tab[TB_CHI * ld + q]; g.xhi = tab[TB_XHI * ld + q];
    g.lprp = tab[TB_LPRP * ld + q]; g.epsz = tab[TB_EPSZ * ld + q];
}

// (an LJ particle's reach is the cutoff)
__device__ __forceinline__ void batch_self(const double *tab, int ld, int q, bool gb, double rc_lj, BatchSelf &me)
{
    me.x = tab[TB_X * ld + q]; me.y = tab[TB_Y * ld + q]; me.z = tab[TB_Z * ld + q];
    me.reach = rc_lj;
    if (gb) {
        batch_gb_self(tab, ld, q, me.g);
        me.reach = tab[TB_REACH * ld + q];
    }
}

// RIJ of the README loop's pair: (q, jq) when jq > q, else (jq, q)
__device__ __forceinline__ void batch_rij(const double *tab, int ld, int jq, bool fwd, const BatchSelf &me, double L,
                                          double r[3])
{
    const double xj = tab[TB_X * ld + jq], yj = tab[TB_Y * ld + jq], zj = tab[TB_Z * ld + jq];
    r[0] = sx_mimg_div(fwd ? me.x - xj : xj - me.x, L);
    r[1] = sx_mimg_div(fwd ? me.y - yj : yj - me.y, L);
    r[2] = sx_mimg_div(fwd ? me.z - zj : zj - me.z, L);
}

// false only where the pair's force, torques and energy are exactly zero:
// beyond the LJ cutoff, or beyond the Gay-Berne reach of the pair's I (an
// FP64 test in both precisions of the pair terms)
__device__ __forceinline__ bool batch_in_reach(const double *tab, int ld, int q, int jq, const BatchSelf &me,
                                               bool gb, double L)
{
    const bool fwd = jq > q;
    double r[3];
    batch_rij(tab, ld, jq, fwd, me, L, r);
    const double rsq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    const double reach = (gb && !fwd) ? tab[TB_REACH * ld + jq] : me.reach;
    return rsq < reach * reach;
}

// adds pair (q, jq)'s contribution to particle q: force, torque, and the
// pair energy once (on the pair's I); T = float: the FP32 pair mode (LJ in
// FP32, gb_ in its mixed form), the sums in FP64 either way
template <typename T>
__device__ __forceinline__ void batch_gb(const double r[3], const GbSelf &A, const double u2[3], const double *p,
                                         int mu, int nu, PairOutT<T> &o)
{
    if constexpr (sizeof(T) == sizeof(double)) gb_core<0>(r, A, u2, p[0], mu, nu, o);
    else gb_core_f32(r, A, u2, p[0], mu, nu, o);
}

template <typename T>
__device__ __forceinline__ void batch_eval(const double *tab, int ld, int q, int jq, const BatchSelf &me, bool gb,
                                           const double *p, int mu, int nu, double L, double acc[7])
{
    const bool fwd = jq > q;
    double r[3];
    batch_rij(tab, ld, jq, fwd, me, L, r);
    PairOutT<T> o;
    if (!gb) {
        lj_core<T>(r, p, o);
        if (fwd) {
            acc[0] = acc[0] + o.f[0]; acc[1] = acc[1] + o.f[1]; acc[2] = acc[2] + o.f[2];
            acc[6] = acc[6] + o.pe;
        } else {
            acc[0] = acc[0] - o.f[0]; acc[1] = acc[1] - o.f[1]; acc[2] = acc[2] - o.f[2];
        }
        return;
    }
    const double uj[3] = {tab[TB_U0 * ld + jq], tab[TB_U1 * ld + jq], tab[TB_U2 * ld + jq]};
    if (fwd) {
        batch_gb(r, me.g, uj, p, mu, nu, o);
        acc[0] = acc[0] + o.f[0]; acc[1] = acc[1] + o.f[1]; acc[2] = acc[2] + o.f[2];
        acc[3] = acc[3] + o.ti[0]; acc[4] = acc[4] + o.ti[1]; acc[5] = acc[5] + o.ti[2];
        acc[6] = acc[6] + o.pe;
    } else {
        GbSelf gj;
        batch_gb_self(tab, ld, jq, gj);
        batch_gb(r, gj, me.g.u, p, mu, nu, o);
        acc[0] = acc[0] - o.f[0]; acc[1] = acc[1] - o.f[1]; acc[2] = acc[2] - o.f[2];
        acc[3] = acc[3] + o.tj[0]; acc[4] = acc[4] + o.tj[1]; acc[5] = acc[5] + o.tj[2];
    }
}

// lgv_initial_step of particle i (Philox counter: its 1-based index in its
// system) and its table row
__global__ void __launch_bounds__(256) k_batch_initial(DState s, const BatchSys *sys, const int *sys_of, double dt,
                                                       double damp, const int *dstep, double *tab)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    const int sid = sys_of[i];
    const BatchSys S = sys[sid];
    DState ls = s;
    ls.box[0] = ls.box[1] = ls.box[2] = S.L;
    ls.periodic = 1;
    LgvConst k;
    k.dt = dt; k.damp = damp; k.T = S.T; k.e1 = S.e1; k.e2 = S.e2; k.kk1 = S.kk1; k.kk2 = S.kk2;
    k.temp_int = S.temp_int; k.save = 0;
    const int q = i - S.off;
    lgv_initial_particle(ls, i, q + 1, k, *dstep);
    double *t = tab + (size_t)TB_N * S.off;
    const int ld = S.n;
    t[TB_X * ld + q] = s.f[SX_RX][i];
    t[TB_Y * ld + q] = s.f[SX_RY][i];
    t[TB_Z * ld + q] = s.f[SX_RZ][i];
    if (S.kind == SX_GB) {
        GbSelf g;
        gb_self(ls, i, (int)S.p[1], g);
        t[TB_U0 * ld + q] = g.u[0];
        t[TB_U1 * ld + q] = g.u[1];
        t[TB_U2 * ld + q] = g.u[2];
        t[TB_CHI * ld + q] = g.chi;
        t[TB_XHI * ld + q] = g.xhi;
        t[TB_LPRP * ld + q] = g.lprp;
        t[TB_EPSZ * ld + q] = g.epsz;
        t[TB_REACH * ld + q] = (fmax(s.f[SX_SIG_X][i], s.f[SX_SIG_Z][i]) + S.p[0] - 1.0) * (1.0 + 1e-9);
    }
}

// initialize_forces + the pair loop for particles [q0, q1) of one system
template <int BT, int LPP, bool LDS, typename T>
__global__ void __launch_bounds__(BT) k_batch_forces(DState s, const BatchSys *sys, const int4 *blocks,
                                                     const double *gtab, double *pe)
{
    extern __shared__ double lds_tab[];
    __shared__ int queue[LPP == 64 ? BT / 64 : 1][128];   // per-wavefront pair queue (LPP = 64)
    const int4 blk = blocks[blockIdx.x];
    const BatchSys S = sys[blk.x];
    const int q0 = blk.y, q1 = blk.z;
    const bool gb = S.kind == SX_GB;
    const int mu = (int)S.p[1], nu = (int)S.p[2];
    const double L = S.L;
    const int ld = S.n;
    const double *src = gtab + (size_t)TB_N * S.off;
    const double *tab = src;
    if (LDS) {
        const int rows = gb ? TB_N : 3;
        for (int e = threadIdx.x; e < rows * ld; e += BT) lds_tab[e] = src[e];
        __syncthreads();
        tab = lds_tab;
    }
    if constexpr (LPP < 64) {
        // LPP lanes per particle; with LPP = 1 a particle's terms arrive in the
        // README loop's order (pairs (J, q), J < q, then (q, J), J > q)
        const int lane = threadIdx.x % LPP, grp = threadIdx.x / LPP;
        for (int q = q0 + grp; q < q1; q += BT / LPP) {
            const int i = S.off + q;
            BatchSelf me;
            batch_self(tab, ld, q, gb, S.p[2], me);
            double acc[7] = {0, 0, 0, 0, 0, 0, 0};
            for (int jq = lane; jq < S.n; jq += LPP) {
                if (jq == q || !batch_in_reach(tab, ld, q, jq, me, gb, L)) continue;
                batch_eval<T>(tab, ld, q, jq, me, gb, S.p, mu, nu, L, acc);
            }
#pragma unroll
            for (int off = LPP / 2; off > 0; off >>= 1)
#pragma unroll
                for (int c = 0; c < 7; ++c) acc[c] = acc[c] + __shfl_xor(acc[c], off, LPP);
            if (lane == 0) {
                s.f[SX_FX][i] = acc[0]; s.f[SX_FY][i] = acc[1]; s.f[SX_FZ][i] = acc[2];
                s.f[SX_TX][i] = acc[3]; s.f[SX_TY][i] = acc[4]; s.f[SX_TZ][i] = acc[5];
                pe[i] = acc[6];
            }
        }
    } else {
        // One wavefront per particle with compaction: the 64 lanes test 64
        // candidates at a time, the ones in reach are appended (ballot +
        // prefix count) to the wavefront's queue in LDS, and every time the
        // queue holds 64 pairs all lanes evaluate one each.  Lanes never idle
        // on pairs beyond the cutoff.
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        int *jq_buf = queue[wv];
        for (int q = q0 + wv; q < q1; q += BT / 64) {
            const int i = S.off + q;
            BatchSelf me;
            batch_self(tab, ld, q, gb, S.p[2], me);
            double acc[7] = {0, 0, 0, 0, 0, 0, 0};
            int cnt = 0;
            for (int base = 0; base < S.n; base += 64) {
                const int jq = base + lane;
                const bool in = jq < S.n && jq != q && batch_in_reach(tab, ld, q, jq, me, gb, L);
                const unsigned long long m = __ballot(in);
                const int pre = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                if (in) jq_buf[cnt + pre] = jq;
                cnt += __popcll(m);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if (cnt >= 64) {
                    const int j = jq_buf[lane];
                    const int carry = cnt - 64 > lane ? jq_buf[64 + lane] : 0;
                    batch_eval<T>(tab, ld, q, j, me, gb, S.p, mu, nu, L, acc);
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    if (cnt - 64 > lane) jq_buf[lane] = carry;
                    cnt -= 64;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
            }
            if (lane < cnt) batch_eval<T>(tab, ld, q, jq_buf[lane], me, gb, S.p, mu, nu, L, acc);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int off = 32; off > 0; off >>= 1)
#pragma unroll
                for (int c = 0; c < 7; ++c) acc[c] = acc[c] + __shfl_xor(acc[c], off, 64);
            if (lane == 0) {
                s.f[SX_FX][i] = acc[0]; s.f[SX_FY][i] = acc[1]; s.f[SX_FZ][i] = acc[2];
                s.f[SX_TX][i] = acc[3]; s.f[SX_TY][i] = acc[4]; s.f[SX_TZ][i] = acc[5];
                pe[i] = acc[6];
            }
        }
    }
}

// lgv_final_step of one system and its thermo: KE, rotational KE, PE (the
// particles' pair energies summed in a fixed order) and T; the last
// workgroup to finish advances the step counter
__global__ void __launch_bounds__(256) k_batch_final(DState s, const BatchSys *sys, double dt, const double *pe,
                                                     double *out, unsigned *ticket, int *dstep, int nsys)
{
    __shared__ double red[256][3];
    const int sid = blockIdx.x;
    const BatchSys S = sys[sid];
    DState ls = s;
    ls.box[0] = ls.box[1] = ls.box[2] = S.L;
    ls.periodic = 1;
    double ke = 0.0, kr = 0.0, pl = 0.0;
    for (int q = threadIdx.x; q < S.n; q += 256) {
        const int i = S.off + q;
        half_kick(ls, i, dt);
        ang_kick(ls, i, dt);
        ke += kin_trans(ls, i);
        kr += kin_rot(ls, i);
        pl += pe[i];
    }
    red[threadIdx.x][0] = ke; red[threadIdx.x][1] = kr; red[threadIdx.x][2] = pl;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int c = 0; c < 3; ++c) red[threadIdx.x][c] = red[threadIdx.x][c] + red[threadIdx.x + o][c];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[4 * sid + 0] = red[0][0];
        out[4 * sid + 1] = red[0][1];
        out[4 * sid + 2] = red[0][2];
        out[4 * sid + 3] = 2.0 * (red[0][0] + red[0][1]) / (double)(6 * S.n - 3);
        // every system's initial step has read the counter (an earlier launch)
        __threadfence();
        if (atomicAdd(ticket, 1u) == (unsigned)nsys - 1) {
            *ticket = 0u;
            *dstep = *dstep + 1;
        }
    }
}

// The whole step in one launch: a grid of workgroups takes whole systems
// from a queue counter, the most expensive first (order), and runs each
// system's lgv_initial_step with its table (in LDS when it fits), its pair
// loop (one compacting wavefront per particle, as k_batch_forces with 64
// lanes), lgv_final_step and its thermo.  A system never spans workgroups,
// so only workgroup barriers occur and every workgroup leaves once the queue
// is empty; the last one to leave resets the queue and advances the step
// counter.  qctr: [0] queue, [1] workgroups done.
template <int BT, typename T>
__global__ void __launch_bounds__(BT) k_batch_step(DState s, const BatchSys *sys, const int *order, int nsys,
                                                   double dt, double damp, int *dstep, double *gtab, double *pe,
                                                   double *out, unsigned *qctr, int lds_ok)
{
    extern __shared__ double lds_tab[];
    __shared__ int queue[BT / 64][128];
    __shared__ double red[BT][3];
    __shared__ int cur;
    const int step = *dstep;   // no workgroup advances it before every one has left
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (;;) {
        if (threadIdx.x == 0) cur = (int)atomicAdd(&qctr[0], 1u);
        __syncthreads();
        const int k = cur;
        __syncthreads();
        if (k >= nsys) break;
        const int sid = order[k];
        const BatchSys S = sys[sid];
        DState ls = s;
        ls.box[0] = ls.box[1] = ls.box[2] = S.L;
        ls.periodic = 1;
        const bool gb = S.kind == SX_GB;
        const int mu = (int)S.p[1], nu = (int)S.p[2];
        const int ld = S.n;
        double *tab = lds_ok ? lds_tab : gtab + (size_t)TB_N * S.off;
        // lgv_initial_step of each particle (counter: its 1-based index in
        // the system) and its table row
        LgvConst kc;
        kc.dt = dt; kc.damp = damp; kc.T = S.T; kc.e1 = S.e1; kc.e2 = S.e2; kc.kk1 = S.kk1; kc.kk2 = S.kk2;
        kc.temp_int = S.temp_int; kc.save = 0;
        for (int q = threadIdx.x; q < S.n; q += BT) {
            const int i = S.off + q;
            lgv_initial_particle(ls, i, q + 1, kc, step);
            tab[TB_X * ld + q] = s.f[SX_RX][i];
            tab[TB_Y * ld + q] = s.f[SX_RY][i];
            tab[TB_Z * ld + q] = s.f[SX_RZ][i];
            if (gb) {
                GbSelf g;
                gb_self(ls, i, mu, g);
                tab[TB_U0 * ld + q] = g.u[0];
                tab[TB_U1 * ld + q] = g.u[1];
                tab[TB_U2 * ld + q] = g.u[2];
                tab[TB_CHI * ld + q] = g.chi;
                tab[TB_XHI * ld + q] = g.xhi;
                tab[TB_LPRP * ld + q] = g.lprp;
                tab[TB_EPSZ * ld + q] = g.epsz;
                tab[TB_REACH * ld + q] = (fmax(s.f[SX_SIG_X][i], s.f[SX_SIG_Z][i]) + S.p[0] - 1.0) * (1.0 + 1e-9);
            }
        }
        __syncthreads();
        // initialize_forces + the pair loop: one compacting wavefront per particle
        int *jq_buf = queue[wv];
        for (int q = wv; q < S.n; q += BT / 64) {
            const int i = S.off + q;
            BatchSelf me;
            batch_self(tab, ld, q, gb, S.p[2], me);
            double acc[7] = {0, 0, 0, 0, 0, 0, 0};
            int cnt = 0;
            for (int base = 0; base < S.n; base += 64) {
                const int jq = base + lane;
                const bool in = jq < S.n && jq != q && batch_in_reach(tab, ld, q, jq, me, gb, S.L);
                const unsigned long long m = __ballot(in);
                const int pre = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                if (in) jq_buf[cnt + pre] = jq;
                cnt += __popcll(m);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if (cnt >= 64) {
                    const int j = jq_buf[lane];
                    const int carry = cnt - 64 > lane ? jq_buf[64 + lane] : 0;
                    batch_eval<T>(tab, ld, q, j, me, gb, S.p, mu, nu, S.L, acc);
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    if (cnt - 64 > lane) jq_buf[lane] = carry;
                    cnt -= 64;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
            }
            if (lane < cnt) batch_eval<T>(tab, ld, q, jq_buf[lane], me, gb, S.p, mu, nu, S.L, acc);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int off = 32; off > 0; off >>= 1)
#pragma unroll
                for (int c = 0; c < 7; ++c) acc[c] = acc[c] + __shfl_xor(acc[c], off, 64);
            if (lane == 0) {
                s.f[SX_FX][i] = acc[0]; s.f[SX_FY][i] = acc[1]; s.f[SX_FZ][i] = acc[2];
                s.f[SX_TX][i] = acc[3]; s.f[SX_TY][i] = acc[4]; s.f[SX_TZ][i] = acc[5];
                pe[i] = acc[6];
            }
        }
        __syncthreads();
        // lgv_final_step and the system's thermo (as k_batch_final)
        double ke = 0.0, kr = 0.0, pl = 0.0;
        for (int q = threadIdx.x; q < S.n; q += BT) {
            const int i = S.off + q;
            half_kick(ls, i, dt);
            ang_kick(ls, i, dt);
            ke += kin_trans(ls, i);
            kr += kin_rot(ls, i);
            pl += pe[i];
        }
        red[threadIdx.x][0] = ke; red[threadIdx.x][1] = kr; red[threadIdx.x][2] = pl;
        __syncthreads();
        for (int o = BT / 2; o > 0; o >>= 1) {
            if (threadIdx.x < o)
                for (int c = 0; c < 3; ++c) red[threadIdx.x][c] = red[threadIdx.x][c] + red[threadIdx.x + o][c];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            out[4 * sid + 0] = red[0][0];
            out[4 * sid + 1] = red[0][1];
            out[4 * sid + 2] = red[0][2];
            out[4 * sid + 3] = 2.0 * (red[0][0] + red[0][1]) / (double)(6 * S.n - 3);
        }
        __syncthreads();   // the table and red are reused by the next system
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&qctr[1], 1u) == gridDim.x - 1) {
            atomicExch(&qctr[0], 0u);
            atomicExch(&qctr[1], 0u);
            *dstep = step + 1;
        }
    }
}

__global__ void k_batch_step_set(int *d, int v) { *d = v; }

static BatchState *batch_of(sx_ctx *c)
{
    if (!c->batch) throw sx_error{SX_ERR_STATE, "no batch defined (sx_batch_set)"};
    BatchState *B = (BatchState *)c->batch;
    if (B->n_total != c->n || B->fields != c->d_fields)
        throw sx_error{SX_ERR_STATE, "particle storage changed since sx_batch_set"};
    return B;
}

void sx_batch_free(sx_ctx *c)
{
    BatchState *B = (BatchState *)c->batch;
    if (!B) return;
    if (B->exec) hipGraphExecDestroy(B->exec);
    if (B->graph) hipGraphDestroy(B->graph);
    for (void *q : {(void *)B->d_sys, (void *)B->d_sys_of, (void *)B->d_tab, (void *)B->d_pe, (void *)B->d_blocks,
                    (void *)B->d_out, (void *)B->d_ticket, (void *)B->d_step, (void *)B->d_order,
                    (void *)B->d_qctr})
        if (q) hipFree(q);
    delete B;
    c->batch = nullptr;
}

static void batch_constants(sx_ctx *c, BatchState *B, double dt, double damp)
{
    if (dt == B->dt && damp == B->damp) return;
    for (auto &S : B->h_sys) {
        S.e1 = std::exp(-damp * dt);
        S.e2 = 1 - std::exp(-2.0 * damp * dt);
        S.kk1 = 1.0 - (damp * dt) / 2.0;
        S.kk2 = 1.0 / (1.0 + (damp * dt) / 2.0);
    }
    HIPCHK(hipMemcpyAsync(B->d_sys, B->h_sys.data(), sizeof(BatchSys) * B->nsys, hipMemcpyHostToDevice, c->stream));
    B->dt = dt;
    B->damp = damp;
}

// force workgroups of a shape: slices of P particles, the most expensive
// first (work ~ slice length x N_s, Gay-Berne weighted 4x)
static void batch_blocks(sx_ctx *c, BatchState *B, int choice)
{
    if (B->blocks_choice == choice || shape_fused(choice)) return;
    const int P = BATCH_SHAPES[choice][0];
    std::vector<int4> blk;
    std::vector<double> cost;
    for (int q = 0; q < B->nsys; ++q) {
        const BatchSys &S = B->h_sys[q];
        for (int a = 0; a < S.n; a += P) {
            int4 b;
            b.x = q; b.y = a; b.z = std::min(S.n, a + P); b.w = 0;
            blk.push_back(b);
            cost.push_back((double)(b.z - b.y) * S.n * (S.kind == SX_GB ? 4.0 : 1.0));
        }
    }
    std::vector<int> ord(blk.size());
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<int4> sorted(blk.size());
    for (size_t k = 0; k < ord.size(); ++k) sorted[k] = blk[ord[k]];
    if ((int)sorted.size() > B->blocks_cap) {
        if (B->d_blocks) HIPCHK(hipFree(B->d_blocks));
        HIPCHK(hipMalloc(&B->d_blocks, sizeof(int4) * sorted.size()));
        B->blocks_cap = (int)sorted.size();
        B->graph_choice = -1;   // a captured step would use the old buffer
    }
    HIPCHK(hipMemcpyAsync(B->d_blocks, sorted.data(), sizeof(int4) * sorted.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    B->nblocks = (int)sorted.size();
    B->blocks_choice = choice;
}

template <typename T>
static void batch_launch_t(sx_ctx *c, BatchState *B, int choice, double dt, double damp)
{
    DState s = c->dstate();
    if (shape_fused(choice)) {
        // the table in LDS when it fits beside the kernel's own arrays (queue,
        // reduction) in 64 KiB per workgroup; else each system's global table
        const int BT = BATCH_SHAPES[choice][1];
        const size_t fixed = (size_t)(BT / 64) * 128 * sizeof(int) + (size_t)BT * 3 * sizeof(double) + 64;
        const size_t tbytes = sizeof(double) * TB_N * (size_t)B->cap;
        const bool lds = tbytes + fixed <= (size_t)BATCH_LDS_BYTES;
        const size_t lbytes = lds ? tbytes : 0;
        const dim3 g(std::min(B->ngrid, B->nsys));
        if (BATCH_SHAPES[choice][1] == 256)
            hipLaunchKernelGGL((k_batch_step<256, T>), g, dim3(256), lbytes, c->stream, s, B->d_sys, B->d_order, B->nsys,
                               dt, damp, B->d_step, B->d_tab, B->d_pe, B->d_out, B->d_qctr, lds ? 1 : 0);
        else
            hipLaunchKernelGGL((k_batch_step<512, T>), g, dim3(512), lbytes, c->stream, s, B->d_sys, B->d_order, B->nsys,
                               dt, damp, B->d_step, B->d_tab, B->d_pe, B->d_out, B->d_qctr, lds ? 1 : 0);
        HIPCHK(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(k_batch_initial, dim3(sx_blocks(c->n, 256)), dim3(256), 0, c->stream, s, B->d_sys,
                       B->d_sys_of, dt, damp, B->d_step, B->d_tab);
    const bool lds = sizeof(double) * TB_N * (size_t)B->cap <= (size_t)BATCH_LDS_BYTES;
    const size_t lbytes = lds ? sizeof(double) * TB_N * (size_t)B->cap : 0;
    dim3 g(B->nblocks);
#define SX_BATCH_CASE(K, BT, LPP)                                                                                  \
    case K:                                                                                                        \
        if (lds)                                                                                                   \
            hipLaunchKernelGGL((k_batch_forces<BT, LPP, true, T>), g, dim3(BT), lbytes, c->stream, s, B->d_sys,       \
                               B->d_blocks, B->d_tab, B->d_pe);                                                    \
        else                                                                                                       \
            hipLaunchKernelGGL((k_batch_forces<BT, LPP, false, T>), g, dim3(BT), 0, c->stream, s, B->d_sys,           \
                               B->d_blocks, B->d_tab, B->d_pe);                                                    \
        break;
    switch (choice) {
        SX_BATCH_CASE(0, 64, 1)
        SX_BATCH_CASE(1, 64, 4)
        SX_BATCH_CASE(2, 128, 4)
        SX_BATCH_CASE(3, 256, 8)
        SX_BATCH_CASE(4, 256, 8)
        SX_BATCH_CASE(5, 256, 64)
        SX_BATCH_CASE(6, 256, 64)
        SX_BATCH_CASE(7, 256, 64)
    default: throw sx_error{SX_ERR_ARG, "bad launch shape"};
    }
#undef SX_BATCH_CASE
    hipLaunchKernelGGL(k_batch_final, dim3(B->nsys), dim3(256), 0, c->stream, s, B->d_sys, dt, B->d_pe, B->d_out,
                       B->d_ticket, B->d_step, B->nsys);
}

// the pair terms in FP32 when the context's pair precision is 32 bits
static void batch_launch(sx_ctx *c, BatchState *B, int choice, double dt, double damp)
{
    if (c->pair_f32) batch_launch_t<float>(c, B, choice, dt, damp);
    else batch_launch_t<double>(c, B, choice, dt, damp);
}

extern "C" {

int sx_batch_set(sx_ctx *c, int nsys, const int *counts, const double *box, const int *kind, const double *params,
                 const double *temp)
{
    return guard(c, [&] {
        need_system(c);
        if (nsys <= 0 || !counts || !box || !kind || !params || !temp) throw sx_error{SX_ERR_ARG, "bad batch"};
        long total = 0;
        for (int q = 0; q < nsys; ++q) total += counts[q];
        if (total != c->n) throw sx_error{SX_ERR_ARG, "system sizes do not add up to the particle count"};
        std::vector<double> sigz(c->n), sigx(c->n);
        HIPCHK(hipMemcpy(sigz.data(), c->d_fields + (size_t)SX_SIG_Z * c->ld, sizeof(double) * c->n,
                         hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(sigx.data(), c->d_fields + (size_t)SX_SIG_X * c->ld, sizeof(double) * c->n,
                         hipMemcpyDeviceToHost));
        sx_batch_free(c);
        struct Owner {
            sx_ctx *c;
            BatchState *b;
            ~Owner()
            {
                if (b) {   // a failed set-up: release what was allocated
                    c->batch = b;
                    sx_batch_free(c);
                }
            }
            BatchState *get() { return b; }
            BatchState *release()
            {
                BatchState *r = b;
                b = nullptr;
                return r;
            }
        } owner{c, new BatchState()};
        BatchState *B = owner.get();
        B->nsys = nsys;
        B->n_total = c->n;
        B->fields = c->d_fields;
        B->h_sys.resize(nsys);
        int off = 0;
        for (int q = 0; q < nsys; ++q) {
            BatchSys &S = B->h_sys[q];
            if (counts[q] < 2) throw sx_error{SX_ERR_ARG, "a system needs at least two particles"};
            if (kind[q] != SX_LJ && kind[q] != SX_GB) throw sx_error{SX_ERR_ARG, "batched kinds are SX_LJ and SX_GB"};
            S.off = off;
            S.n = counts[q];
            S.kind = kind[q];
            S.L = box[q];
            S.T = temp[q];
            S.temp_int = (int)(S.T * 1000.0);   // integrators.f90:532 (host FP64)
            for (int k = 0; k < 4; ++k) S.p[k] = params[4 * q + k];
            double smax = 0.0;
            for (int a = off; a < off + counts[q]; ++a) smax = std::max(smax, std::max(sigz[a], sigx[a]));
            // the minimum image must see every interacting pair (system.f90:902-906)
            double range = S.kind == SX_LJ ? S.p[2] : smax + S.p[0] - 1.0;
            if (range > 0.5 * S.L)
                throw sx_error{SX_ERR_CUTOFF, "interaction range exceeds half the box of system " + std::to_string(q)};
            off += counts[q];
        }
        B->cap = *std::max_element(counts, counts + nsys);
        std::vector<int> sys_of(c->n);
        for (int q = 0; q < nsys; ++q)
            for (int a = 0; a < counts[q]; ++a) sys_of[B->h_sys[q].off + a] = q;
        HIPCHK(hipMalloc(&B->d_sys, sizeof(BatchSys) * nsys));
        HIPCHK(hipMalloc(&B->d_sys_of, sizeof(int) * c->n));
        HIPCHK(hipMalloc(&B->d_tab, sizeof(double) * TB_N * (size_t)c->n));
        HIPCHK(hipMalloc(&B->d_pe, sizeof(double) * (size_t)c->n));
        HIPCHK(hipMemcpyAsync(B->d_sys_of, sys_of.data(), sizeof(int) * c->n, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMalloc(&B->d_out, sizeof(double) * 4 * nsys));
        HIPCHK(hipMalloc(&B->d_ticket, sizeof(unsigned)));
        HIPCHK(hipMalloc(&B->d_step, sizeof(int)));
        HIPCHK(hipMemsetAsync(B->d_out, 0, sizeof(double) * 4 * nsys, c->stream));
        HIPCHK(hipMemsetAsync(B->d_ticket, 0, sizeof(unsigned), c->stream));
        // the one-launch step takes systems in this order: pair work ~ N^2,
        // Gay-Berne weighted 4x, largest first
        std::vector<int> order(nsys);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            auto w = [&](int q) {
                return (double)B->h_sys[q].n * B->h_sys[q].n * (B->h_sys[q].kind == SX_GB ? 4.0 : 1.0);
            };
            return w(a) > w(b);
        });
        HIPCHK(hipMalloc(&B->d_order, sizeof(int) * nsys));
        HIPCHK(hipMemcpyAsync(B->d_order, order.data(), sizeof(int) * nsys, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMalloc(&B->d_qctr, sizeof(unsigned) * 2));
        HIPCHK(hipMemsetAsync(B->d_qctr, 0, sizeof(unsigned) * 2, c->stream));
        int ncu = 0;
        HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
        B->ngrid = std::max(1, 2 * ncu);
        HIPCHK(hipStreamSynchronize(c->stream));
        c->batch = owner.release();
    });
}

// nsteps whole steps of every system; step = first_step, first_step + 1, ...
int sx_batch_run(sx_ctx *c, int nsteps, double dt, double damping, int first_step)
{
    return guard(c, [&] {
        BatchState *B = batch_of(c);
        if (nsteps < 0) throw sx_error{SX_ERR_ARG, "nsteps < 0"};
        batch_constants(c, B, dt, damping);
        batch_blocks(c, B, B->choice);
        hipLaunchKernelGGL(k_batch_step_set, dim3(1), dim3(1), 0, c->stream, B->d_step, first_step);
        if (nsteps == 0) return;
        const char *ng = getenv("SUCROSE_NO_GRAPH");
        if (ng && ng[0] == '1') {
            for (int t = 0; t < nsteps; ++t) batch_launch(c, B, B->choice, dt, damping);
            sx_check_device_error(c);
            return;
        }
        if (!B->exec || B->graph_choice != B->choice || B->graph_dt != dt || B->graph_damp != damping ||
            B->graph_f32 != c->pair_f32) {
            if (B->exec) { hipGraphExecDestroy(B->exec); B->exec = nullptr; }
            if (B->graph) { hipGraphDestroy(B->graph); B->graph = nullptr; }
            HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            batch_launch(c, B, B->choice, dt, damping);
            HIPCHK(hipStreamEndCapture(c->stream, &B->graph));
            HIPCHK(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
            B->graph_choice = B->choice;
            B->graph_dt = dt;
            B->graph_damp = damping;
            B->graph_f32 = c->pair_f32;
        }
        for (int t = 0; t < nsteps; ++t) HIPCHK(hipGraphLaunch(B->exec, c->stream));
        sx_check_device_error(c);
    });
}

// per system: kinetic, rotational kinetic and potential energy and the
// temperature 2 (KE + KErot) / (6 N - 3) after the last step; out[4 nsys]
int sx_batch_thermo(sx_ctx *c, double *out)
{
    return guard(c, [&] {
        BatchState *B = batch_of(c);
        HIPCHK(hipMemcpyAsync(out, B->d_out, sizeof(double) * 4 * B->nsys, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

// Times nrep steps of every launch shape from the same saved state (which is
// restored afterwards) and keeps the fastest; ms[8] (may be NULL) receives the
// time per step of each shape, *choice the chosen index.
int sx_batch_autotune(sx_ctx *c, double dt, double damping, int first_step, int nrep, int *choice, double *ms)
{
    return guard(c, [&] {
        BatchState *B = batch_of(c);
        if (nrep < 1) nrep = 1;
        batch_constants(c, B, dt, damping);
        const size_t bytes = sizeof(double) * (size_t)SX_NFIELDS * (size_t)c->ld;
        double *saved = nullptr;
        HIPCHK(hipMalloc(&saved, bytes));
        HIPCHK(hipMemcpyAsync(saved, c->d_fields, bytes, hipMemcpyDeviceToDevice, c->stream));
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        double best = 1e300;
        int best_k = B->choice;
        for (int k = 0; k < BATCH_NSHAPES; ++k) {
            batch_blocks(c, B, k);
            hipLaunchKernelGGL(k_batch_step_set, dim3(1), dim3(1), 0, c->stream, B->d_step, first_step);
            batch_launch(c, B, k, dt, damping);   // warm
            HIPCHK(hipEventRecord(e0, c->stream));
            for (int r = 0; r < nrep; ++r) batch_launch(c, B, k, dt, damping);
            HIPCHK(hipEventRecord(e1, c->stream));
            HIPCHK(hipEventSynchronize(e1));
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, e0, e1));
            double per = (double)t / nrep;
            if (ms) ms[k] = per;
            if (per < best) { best = per; best_k = k; }
            HIPCHK(hipMemcpyAsync(c->d_fields, saved, bytes, hipMemcpyDeviceToDevice, c->stream));
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(saved));
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        B->choice = best_k;
        if (choice) *choice = best_k;
        sx_check_device_error(c);
    });
}

int sx_batch_set_shape(sx_ctx *c, int choice)
{
    return guard(c, [&] {
        BatchState *B = batch_of(c);
        if (choice < 0 || choice >= BATCH_NSHAPES) throw sx_error{SX_ERR_ARG, "bad launch shape"};
        B->choice = choice;
    });
}

}  // extern "C"
