// Batched replicas: many independent small systems advanced together, one
// workgroup per system and ONE kernel launch per step (captured in a HIP
// graph and replayed).  Each system is a contiguous range of the context's
// particle arrays with its own cubic periodic box, pair potential (LJ or
// Gay-Berne), parameters and Langevin temperature.  Per system and step the
// workgroup runs the reference's loop (README.md:83-115):
// lgv_initial_step, initialize_forces, the pair loop over all pairs I < J of
// the system (the README's own double loop; the cutoff of each potential does
// the screening, so no neighbour list is built), lgv_final_step and
// calculate_thermo's KE / rotational KE / PE / T for that system.
//
// Forces are gathered per particle (LPP lanes per particle, combined with
// cross-lane shuffles in a fixed order): no atomics, deterministic results.
// Systems are dealt to workgroups largest first so the hardware scheduler
// balances the uneven sizes.  The launch shape (workgroup size, lanes per
// particle) is picked by sx_batch_autotune from timed trial steps.
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
    std::vector<int> h_order;
    BatchSys *d_sys = nullptr;
    int *d_order = nullptr;
    double *d_out = nullptr;       // per system: KE, KErot, PE, T
    unsigned *d_ticket = nullptr;
    int *d_step = nullptr;
    int choice = 3;
    double dt = -1, damp = -1;     // constants currently uploaded
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_choice = -1;
    double graph_dt = -1, graph_damp = -1;   // kernel arguments baked into the graph
    int n_total = 0;                         // particle count and field storage at sx_batch_set
    const double *fields = nullptr;
    int cap = 0;                             // largest system
    bool lds = true;                         // per-particle table in LDS (else d_tab)
    double *d_tab = nullptr;
};

// (workgroup size, lanes per particle); 64 lanes = one compacting wavefront per particle
static const int BATCH_SHAPES[][2] = {{64, 1}, {128, 1}, {128, 2}, {256, 2}, {256, 4}, {256, 8}, {256, 64}, {512, 64}};
static const int BATCH_NSHAPES = 8;

// Per-particle table of the system being stepped, filled after
// lgv_initial_step: position, and for Gay-Berne the factors of gb_ that depend
// on one particle only (GbSelf) plus its reach, an upper bound on the distance
// at which the pair with that particle as I can interact:
// max(sig_x, sig_z) + cutoff - sigma_0 bounds gb_'s sigma + cutoff - sigma_0
// because sigma(u1, u2, r) never exceeds the larger of the two diameters.
enum { TB_X, TB_Y, TB_Z, TB_U0, TB_U1, TB_U2, TB_CHI, TB_XHI, TB_LPRP, TB_EPSZ, TB_REACH, TB_N };
static const int BATCH_LDS_BYTES = 64 * 1024;

// particle q of the table: position and (Gay-Berne) its own factors and reach
struct BatchSelf {
    double x, y, z, reach;
    GbSelf g;
};

// (an LJ particle's reach is the cutoff)
__device__ __forceinline__ void batch_self(const double *tab, int ld, int q, bool gb, double rc_lj, BatchSelf &me)
{
    me.x = tab[TB_X * ld + q]; me.y = tab[TB_Y * ld + q]; me.z = tab[TB_Z * ld + q];
    me.reach = rc_lj;
    if (gb) {
        me.g.u[0] = tab[TB_U0 * ld + q]; me.g.u[1] = tab[TB_U1 * ld + q]; me.g.u[2] = tab[TB_U2 * ld + q];
        me.g.chi = tab[TB_CHI * ld + q]; me.g.xhi = tab[TB_XHI * ld + q];
        me.g.lprp = tab[TB_LPRP * ld + q]; me.g.epsz = tab[TB_EPSZ * ld + q];
        me.reach = tab[TB_REACH * ld + q];
    }
}

// RIJ of the README loop's pair: (q, jq) when jq > q, else (jq, q)
__device__ __forceinline__ void batch_rij(const double *tab, int ld, int jq, bool fwd, const BatchSelf &me, double L,
                                          double r[3])
{
    const double xj = tab[TB_X * ld + jq], yj = tab[TB_Y * ld + jq], zj = tab[TB_Z * ld + jq];
    r[0] = sx_mimg(fwd ? me.x - xj : xj - me.x, L);
    r[1] = sx_mimg(fwd ? me.y - yj : yj - me.y, L);
    r[2] = sx_mimg(fwd ? me.z - zj : zj - me.z, L);
}

// false only where the pair's force, torques and energy are exactly zero:
// beyond the LJ cutoff, or beyond the Gay-Berne reach of the pair's I
__device__ __forceinline__ bool batch_in_reach(const double *tab, int ld, int q, int jq, const BatchSelf &me, bool gb,
                                               double L)
{
    const bool fwd = jq > q;
    double r[3];
    batch_rij(tab, ld, jq, fwd, me, L, r);
    const double rsq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    const double reach = (gb && !fwd) ? tab[TB_REACH * ld + jq] : me.reach;
    return rsq < reach * reach;
}

// adds pair (q, jq)'s contribution to particle q: force, torque, and the
// pair energy once (on the pair's I)
__device__ __forceinline__ void batch_eval(const double *tab, int ld, int q, int jq, const BatchSelf &me, bool gb,
                                           const double *p, int mu, int nu, double L, double acc[7])
{
    const bool fwd = jq > q;
    double r[3];
    batch_rij(tab, ld, jq, fwd, me, L, r);
    PairOut o;
    if (!gb) {
        lj_core(r, p, o);
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
        gb_core<0>(r, me.g, uj, p[0], mu, nu, o);
        acc[0] = acc[0] + o.f[0]; acc[1] = acc[1] + o.f[1]; acc[2] = acc[2] + o.f[2];
        acc[3] = acc[3] + o.ti[0]; acc[4] = acc[4] + o.ti[1]; acc[5] = acc[5] + o.ti[2];
        acc[6] = acc[6] + o.pe;
    } else {
        GbSelf gj;
        gj.u[0] = uj[0]; gj.u[1] = uj[1]; gj.u[2] = uj[2];
        gj.chi = tab[TB_CHI * ld + jq]; gj.xhi = tab[TB_XHI * ld + jq];
        gj.lprp = tab[TB_LPRP * ld + jq]; gj.epsz = tab[TB_EPSZ * ld + jq];
        gb_core<0>(r, gj, me.g.u, p[0], mu, nu, o);
        acc[0] = acc[0] - o.f[0]; acc[1] = acc[1] - o.f[1]; acc[2] = acc[2] - o.f[2];
        acc[3] = acc[3] + o.tj[0]; acc[4] = acc[4] + o.tj[1]; acc[5] = acc[5] + o.tj[2];
    }
}

template <int BT, int LPP, bool LDS>
__global__ void __launch_bounds__(BT) k_batch_step(DState s, const BatchSys *sys, const int *order, double dt,
                                                   double damp, int step, int *dstep, double *out, unsigned *ticket,
                                                   int nsys, double *gtab, int cap)
{
    extern __shared__ double lds_tab[];
    __shared__ double red[BT][3];
    __shared__ int queue[LPP == 64 ? BT / 64 : 1][128];   // per-wavefront pair queue (LPP = 64)
    const int sid = order[blockIdx.x];
    const BatchSys S = sys[sid];
    DState ls = s;
    ls.box[0] = ls.box[1] = ls.box[2] = S.L;
    ls.periodic = 1;
    const int st = dstep ? *(volatile const int *)dstep : step;
    LgvConst k;
    k.dt = dt; k.damp = damp; k.T = S.T; k.e1 = S.e1; k.e2 = S.e2; k.kk1 = S.kk1; k.kk2 = S.kk2;
    k.temp_int = S.temp_int; k.save = 0;
    // lgv_initial_step; the Philox counter is the particle's index in its system
    for (int q = threadIdx.x; q < S.n; q += BT) lgv_initial_particle(ls, S.off + q, q + 1, k, st);
    __syncthreads();
    double *tab = LDS ? lds_tab : gtab + (size_t)TB_N * S.off;
    const int ld = LDS ? cap : S.n;
    const bool gb = S.kind == SX_GB;
    const int mu = (int)S.p[1], nu = (int)S.p[2];
    const double L = S.L;
    for (int q = threadIdx.x; q < S.n; q += BT) {
        const int i = S.off + q;
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
    // initialize_forces + the pair loop, gathered per particle
    double pe_loc = 0.0;
    if constexpr (LPP < 64) {
        // LPP lanes per particle; for particle q the contributions arrive in
        // the README loop's order (pairs (J, q), J < q, then (q, J), J > q)
        // when LPP = 1
        const int lane = threadIdx.x % LPP, grp = threadIdx.x / LPP;
        for (int q = grp; q < S.n; q += BT / LPP) {
            const int i = S.off + q;
            BatchSelf me;
            batch_self(tab, ld, q, gb, S.p[2], me);
            double acc[7] = {0, 0, 0, 0, 0, 0, 0};
            for (int jq = lane; jq < S.n; jq += LPP) {
                if (jq == q || !batch_in_reach(tab, ld, q, jq, me, gb, L)) continue;
                batch_eval(tab, ld, q, jq, me, gb, S.p, mu, nu, L, acc);
            }
#pragma unroll
            for (int off = LPP / 2; off > 0; off >>= 1)
#pragma unroll
                for (int c = 0; c < 7; ++c) acc[c] = acc[c] + __shfl_xor(acc[c], off, LPP);
            if (lane == 0) {
                s.f[SX_FX][i] = acc[0]; s.f[SX_FY][i] = acc[1]; s.f[SX_FZ][i] = acc[2];
                s.f[SX_TX][i] = acc[3]; s.f[SX_TY][i] = acc[4]; s.f[SX_TZ][i] = acc[5];
                pe_loc += acc[6];
            }
        }
    } else {
        // One wavefront per particle with compaction: the 64 lanes test 64
        // candidates at a time, the ones in reach are appended (ballot +
        // prefix count) to the wavefront's queue in LDS, and every time the
        // queue holds 64 pairs all lanes evaluate one each.  Lanes never idle
        // on pairs beyond the cutoff.  Sums are combined in a fixed order, so
        // results are deterministic.
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        int *jq_buf = queue[wv];
        for (int q = wv; q < S.n; q += BT / 64) {
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
                    batch_eval(tab, ld, q, j, me, gb, S.p, mu, nu, L, acc);
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
            if (lane < cnt) batch_eval(tab, ld, q, jq_buf[lane], me, gb, S.p, mu, nu, L, acc);
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
                pe_loc += acc[6];
            }
        }
    }
    __syncthreads();
    // lgv_final_step and the system's thermo sums
    double ke = 0.0, kr = 0.0;
    for (int q = threadIdx.x; q < S.n; q += BT) {
        const int i = S.off + q;
        half_kick(ls, i, dt);
        ang_kick(ls, i, dt);
        ke += kin_trans(ls, i);
        kr += kin_rot(ls, i);
    }
    red[threadIdx.x][0] = ke; red[threadIdx.x][1] = kr; red[threadIdx.x][2] = pe_loc;
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
        // the step counter advances once every system has read it
        __threadfence();
        unsigned t = atomicAdd(ticket, 1u);
        if (t == (unsigned)nsys - 1) {
            *ticket = 0u;
            if (dstep) *dstep = st + 1;
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
    for (void *q : {(void *)B->d_sys, (void *)B->d_order, (void *)B->d_out, (void *)B->d_ticket, (void *)B->d_step,
                    (void *)B->d_tab})
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

static void batch_launch(sx_ctx *c, BatchState *B, int choice, double dt, double damp, int step, int *dstep)
{
    dim3 g(B->nsys);
    DState s = c->dstate();
    const size_t lds = B->lds ? sizeof(double) * TB_N * (size_t)B->cap : 0;
#define SX_BATCH_CASE(K, BT, LPP)                                                                                 \
    case K:                                                                                                       \
        if (B->lds)                                                                                               \
            hipLaunchKernelGGL((k_batch_step<BT, LPP, true>), g, dim3(BT), lds, c->stream, s, B->d_sys,          \
                               B->d_order, dt, damp, step, dstep, B->d_out, B->d_ticket, B->nsys, B->d_tab,       \
                               B->cap);                                                                           \
        else                                                                                                      \
            hipLaunchKernelGGL((k_batch_step<BT, LPP, false>), g, dim3(BT), 0, c->stream, s, B->d_sys,           \
                               B->d_order, dt, damp, step, dstep, B->d_out, B->d_ticket, B->nsys, B->d_tab,       \
                               B->cap);                                                                           \
        break;
    switch (choice) {
        SX_BATCH_CASE(0, 64, 1)
        SX_BATCH_CASE(1, 128, 1)
        SX_BATCH_CASE(2, 128, 2)
        SX_BATCH_CASE(3, 256, 2)
        SX_BATCH_CASE(4, 256, 4)
        SX_BATCH_CASE(5, 256, 8)
        SX_BATCH_CASE(6, 256, 64)
        SX_BATCH_CASE(7, 512, 64)
    default: throw sx_error{SX_ERR_ARG, "bad launch shape"};
    }
#undef SX_BATCH_CASE
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
        std::vector<double> sigz(c->n);
        HIPCHK(hipMemcpy(sigz.data(), c->d_fields + (size_t)SX_SIG_Z * c->ld, sizeof(double) * c->n,
                         hipMemcpyDeviceToHost));
        sx_batch_free(c);
        BatchState *B = new BatchState();
        c->batch = B;
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
            double smax = *std::max_element(sigz.begin() + off, sigz.begin() + off + counts[q]);
            // the minimum image must see every interacting pair (system.f90:902-906)
            double range = S.kind == SX_LJ ? S.p[2] : smax + S.p[0] - 1.0;
            if (range > 0.5 * S.L)
                throw sx_error{SX_ERR_CUTOFF, "interaction range exceeds half the box of system " + std::to_string(q)};
            off += counts[q];
        }
        B->cap = *std::max_element(counts, counts + nsys);
        B->lds = sizeof(double) * TB_N * (size_t)B->cap <= (size_t)BATCH_LDS_BYTES;
        if (!B->lds) HIPCHK(hipMalloc(&B->d_tab, sizeof(double) * TB_N * (size_t)c->n));
        B->h_order.resize(nsys);
        std::iota(B->h_order.begin(), B->h_order.end(), 0);
        std::stable_sort(B->h_order.begin(), B->h_order.end(),
                         [&](int a, int b) { return counts[a] > counts[b]; });
        HIPCHK(hipMalloc(&B->d_sys, sizeof(BatchSys) * nsys));
        HIPCHK(hipMalloc(&B->d_order, sizeof(int) * nsys));
        HIPCHK(hipMalloc(&B->d_out, sizeof(double) * 4 * nsys));
        HIPCHK(hipMalloc(&B->d_ticket, sizeof(unsigned)));
        HIPCHK(hipMalloc(&B->d_step, sizeof(int)));
        HIPCHK(hipMemcpyAsync(B->d_order, B->h_order.data(), sizeof(int) * nsys, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemsetAsync(B->d_out, 0, sizeof(double) * 4 * nsys, c->stream));
        HIPCHK(hipMemsetAsync(B->d_ticket, 0, sizeof(unsigned), c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

// nsteps whole steps of every system; step = first_step, first_step + 1, ...
int sx_batch_run(sx_ctx *c, int nsteps, double dt, double damping, int first_step)
{
    return guard(c, [&] {
        BatchState *B = batch_of(c);
        if (nsteps < 0) throw sx_error{SX_ERR_ARG, "nsteps < 0"};
        batch_constants(c, B, dt, damping);
        hipLaunchKernelGGL(k_batch_step_set, dim3(1), dim3(1), 0, c->stream, B->d_step, first_step);
        if (nsteps == 0) return;
        const char *ng = getenv("SUCROSE_NO_GRAPH");
        if (ng && ng[0] == '1') {
            for (int t = 0; t < nsteps; ++t) batch_launch(c, B, B->choice, dt, damping, 0, B->d_step);
            sx_check_device_error(c);
            return;
        }
        if (!B->exec || B->graph_choice != B->choice || B->graph_dt != dt || B->graph_damp != damping) {
            if (B->exec) { hipGraphExecDestroy(B->exec); B->exec = nullptr; }
            if (B->graph) { hipGraphDestroy(B->graph); B->graph = nullptr; }
            HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            batch_launch(c, B, B->choice, dt, damping, 0, B->d_step);
            HIPCHK(hipStreamEndCapture(c->stream, &B->graph));
            HIPCHK(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
            B->graph_choice = B->choice;
            B->graph_dt = dt;
            B->graph_damp = damping;
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
            hipLaunchKernelGGL(k_batch_step_set, dim3(1), dim3(1), 0, c->stream, B->d_step, first_step);
            batch_launch(c, B, k, dt, damping, 0, B->d_step);   // warm
            HIPCHK(hipEventRecord(e0, c->stream));
            for (int r = 0; r < nrep; ++r) batch_launch(c, B, k, dt, damping, 0, B->d_step);
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
