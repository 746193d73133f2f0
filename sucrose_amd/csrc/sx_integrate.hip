// Integrators (integrators.f90), RATTLE (system.f90:1055-1209) and bonded
// terms (force.f90:2244-2462) on the device.  Per-particle updates keep the
// reference's FP64 expressions; loop-invariant transcendental constants
// (exp(-gamma dt), ...) are evaluated once on the host exactly as the
// reference evaluates them and passed in.
#include "sx_ecp.h"
#include "sx_steps.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <numeric>

template <int W>
__device__ __forceinline__ void block_sum_store(double (&v)[W], double *part)
{
    __shared__ double sh[W][256];
    int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < W; ++q) sh[q][t] = v[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
#pragma unroll
            for (int q = 0; q < W; ++q) sh[q][t] = sh[q][t] + sh[q][t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
#pragma unroll
        for (int q = 0; q < W; ++q) part[blockIdx.x * W + q] = sh[q][0];
    }
}

// ---------------------------------------------------------------------------
// velocity Verlet (integrators.f90:7-108)
// ---------------------------------------------------------------------------
__global__ void k_vv_initial(DState s, double dt, int save)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    half_kick(s, i, dt);
    if (save) save_old(s, i);
    s.f[SX_RX][i] = s.f[SX_RX][i] + dt * s.f[SX_VX][i];
    s.f[SX_RY][i] = s.f[SX_RY][i] + dt * s.f[SX_VY][i];
    s.f[SX_RZ][i] = s.f[SX_RZ][i] + dt * s.f[SX_VZ][i];
    wrap_pos(s, i);
}

// calculate_thermo (system.f90:725-741) folded into the final-step kernels
// when the step driver asks for it
struct ThermoArgs {
    int on, nc, nrigid;
    double volume;
    const double *pe_acc;   // non-NULL with `on`: pot_eng = sum of per-particle energies
};

__device__ __forceinline__ void thermo_finish(double *scal, double vir, ThermoArgs th)
{
    double w = vir;
    double kin_tot = scal[SX_KIN_ENG] + scal[SX_KIN_ROT];
    scal[SX_TEMP] = 2.0 * kin_tot / scal[SX_DOF];
    if (th.nc > 0 && scal[SX_VIRIAL_BONDED] >= 0.0) w = w + scal[SX_VIRIAL_BONDED];
    scal[SX_VIRIAL] = w;
    scal[SX_PRESSURE] = (scal[SX_DENSITY] * th.nrigid * scal[SX_TEMP]) + (scal[SX_VIRIAL] / (3.0 * th.volume));
}

__device__ __forceinline__ double virial_of(const DState &s, int i)
{
    return s.f[SX_RX][i] * s.f[SX_FX][i] + s.f[SX_RY][i] * s.f[SX_FY][i] + s.f[SX_RZ][i] * s.f[SX_FZ][i];
}

__global__ void __launch_bounds__(256) k_vv_final(DState s, double dt, double *part, unsigned *ticket, double *scal,
                                                  ThermoArgs th)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    if (i < s.n) {
        half_kick(s, i, dt);
        v[0] = kin_trans(s, i);
        if (th.on) v[1] = virial_of(s, i);
        if (th.pe_acc) v[2] = th.pe_acc[i];
    }
    double tot[3];
    if (!sx_block_reduce_last<3>(v, part, ticket, tot) || threadIdx.x != 0) return;
    scal[SX_KIN_ENG] = tot[0];
    if (th.pe_acc) scal[SX_POT_ENG] = tot[2];
    if (th.on) thermo_finish(scal, tot[1], th);
}

// ---------------------------------------------------------------------------
// quaternion (integrators.f90:123-218)
// ---------------------------------------------------------------------------
__global__ void k_qq_initial(DState s, double dt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    ang_kick(s, i, dt);
    double qdw, qdx, qdy, qdz, RM[9], bx, by, bz;
    quat_midpoint(s, i, dt, qdw, qdx, qdy, qdz, RM, bx, by, bz, true);
    quat_advance(s, i, dt, qdw, qdx, qdy, qdz);
}

__global__ void __launch_bounds__(256) k_qq_final(DState s, double dt, double *part, unsigned *ticket, double *scal,
                                                  ThermoArgs th)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    if (i < s.n) {
        ang_kick(s, i, dt);
        v[0] = kin_rot(s, i);
        if (th.on) v[1] = virial_of(s, i);
        if (th.pe_acc) v[2] = th.pe_acc[i];
    }
    double tot[3];
    if (!sx_block_reduce_last<3>(v, part, ticket, tot) || threadIdx.x != 0) return;
    scal[SX_KIN_ROT] = tot[0];
    if (th.pe_acc) scal[SX_POT_ENG] = tot[2];
    if (th.on) thermo_finish(scal, tot[1], th);
}

// ---------------------------------------------------------------------------
// Langevin (integrators.f90:507-855)
// ---------------------------------------------------------------------------
__global__ void k_lgv_initial(DState s, LgvConst k, int step, const int *dstep)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    int st = dstep ? *dstep : step;
    lgv_initial_particle(s, i, i + 1, k, st);
}

// The start of a whole Langevin step in one launch: lgv_initial_step
// (integrators.f90:507-627), the rebuild decision of needs_rebuild
// (system.f90:1005-1039: maximum displacement since the last build, taken by
// the last workgroup to finish) and initialize_forces (system.f90:423-436:
// each particle's F, T and energy slot after its half kick has read F); with
// PP also the ECP per-particle records of the new orientations.
// the particles' moves since the list was built, |mimg(r - r_last)|^2
__device__ __forceinline__ double moved_sq(const DState &s, int i, const double *xl, const double *yl,
                                           const double *zl)
{
    double dx = s.f[SX_RX][i] - xl[i], dy = s.f[SX_RY][i] - yl[i], dz = s.f[SX_RZ][i] - zl[i];
    if (s.periodic) {
        dx = sx_mimg(dx, s.box[0]); dy = sx_mimg(dy, s.box[1]); dz = sx_mimg(dz, s.box[2]);
    }
    return dx * dx + dy * dy + dz * dz;
}

// needs_rebuild (system.f90:1005-1039) from each thread's move m: per-block
// maxima, the last block to finish takes the maximum over blocks and writes
// the rebuild flag (max is order-free); every thread of the block calls it
template <int BS>
__device__ __forceinline__ void rebuild_flag_from(double m, double *part, unsigned *ticket, double thresh, int rule,
                                                  int *flag)
{
    __shared__ double sh[BS];
    __shared__ bool last;
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int o = BS / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double p[1] = {sh[0]};
        last = sx_publish_take_ticket<1>(part, p, ticket);
    }
    __syncthreads();
    if (!last) return;
    double v = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += BS)
        v = fmax(v, __hip_atomic_load(&part[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = BS / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double mx = sh[0];
        flag[0] = rule == SX_REBUILD_REFERENCE ? (4.0 * mx > thresh) : (mx > thresh);
        sx_ticket_reset(ticket);
    }
}

template <bool PP, int BS>
__global__ void __launch_bounds__(BS) k_lgv_step_begin(DState s, LgvConst k, const int *dstep, const double *xl,
                                                        const double *yl, const double *zl, double *part,
                                                        unsigned *ticket, double thresh, int rule, int *flag,
                                                        double *scal, double *pe_acc, double *pp, float *scr)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (i == 0) {
        scal[SX_POT_ENG] = 0.0; scal[SX_KIN_ENG] = 0.0; scal[SX_KIN_ROT] = 0.0; scal[SX_TEMP] = 0.0;
        scal[SX_PRESSURE] = 0.0; scal[SX_VIRIAL] = 0.0; scal[SX_VIRIAL_BONDED] = 0.0;
    }
    if (i < s.n) {
        lgv_initial_particle(s, i, i + 1, k, *dstep);
        m = moved_sq(s, i, xl, yl, zl);
        s.f[SX_FX][i] = 0.0; s.f[SX_FY][i] = 0.0; s.f[SX_FZ][i] = 0.0;
        s.f[SX_TX][i] = 0.0; s.f[SX_TY][i] = 0.0; s.f[SX_TZ][i] = 0.0;
        pe_acc[i] = 0.0;
        if (PP) ecp_pp_particle(s, i, pp, scr);
    }
    rebuild_flag_from<BS>(m, part, ticket, thresh, rule, flag);
}

// qq_initial_step (integrators.f90:123-171) of each particle, the rebuild
// decision and initialize_forces in one launch: the Nose-Hoover step of the
// step driver, whose chain kernels have already kicked and moved the
// particles (F and T are cleared after the angular kick has read T)
__global__ void __launch_bounds__(256) k_qq_step_begin(DState s, double dt, const double *xl, const double *yl,
                                                       const double *zl, double *part, unsigned *ticket,
                                                       double thresh, int rule, int *flag, double *scal,
                                                       double *pe_acc)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (i == 0) {
        scal[SX_POT_ENG] = 0.0; scal[SX_KIN_ENG] = 0.0; scal[SX_KIN_ROT] = 0.0; scal[SX_TEMP] = 0.0;
        scal[SX_PRESSURE] = 0.0; scal[SX_VIRIAL] = 0.0; scal[SX_VIRIAL_BONDED] = 0.0;
    }
    if (i < s.n) {
        ang_kick(s, i, dt);
        double qdw, qdx, qdy, qdz, RM[9], bx, by, bz;
        quat_midpoint(s, i, dt, qdw, qdx, qdy, qdz, RM, bx, by, bz, true);
        quat_advance(s, i, dt, qdw, qdx, qdy, qdz);
        m = moved_sq(s, i, xl, yl, zl);
        s.f[SX_FX][i] = 0.0; s.f[SX_FY][i] = 0.0; s.f[SX_FZ][i] = 0.0;
        s.f[SX_TX][i] = 0.0; s.f[SX_TY][i] = 0.0; s.f[SX_TZ][i] = 0.0;
        pe_acc[i] = 0.0;
    }
    rebuild_flag_from<256>(m, part, ticket, thresh, rule, flag);
}

// APPLY: the ECP pair results are added to the particle's forces, torques
// and energy first (the apply stage, ecp_particle_sums), and the kicks use
// the sums from registers
template <int BS, bool APPLY>
__global__ void __launch_bounds__(BS) k_lgv_final(DState s, double dt, double *part, unsigned *ticket, double *scal,
                                                  ThermoArgs th, int *dstep, EcpApplyArgs ap,
                                                  const unsigned char *live, double *pe_acc)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < s.n) {
        if (APPLY) {
            // the particle's state, requested before the pair sums so that
            // its latency overlaps theirs
            double f[7] = {s.f[SX_FX][i], s.f[SX_FY][i], s.f[SX_FZ][i], s.f[SX_TX][i],
                           s.f[SX_TY][i], s.f[SX_TZ][i], pe_acc[i]};
            const double im = s.f[SX_INV_MASS][i], mass = s.f[SX_MASS][i];
            double vx = s.f[SX_VX][i], vy = s.f[SX_VY][i], vz = s.f[SX_VZ][i];
            double lx = s.f[SX_LX][i], ly = s.f[SX_LY][i], lz = s.f[SX_LZ][i];
            const double q[4] = {s.f[SX_QW][i], s.f[SX_QX][i], s.f[SX_QY][i], s.f[SX_QZ][i]};
            const double ii[3] = {s.f[SX_IXX_INV][i], s.f[SX_IYY_INV][i], s.f[SX_IZZ_INV][i]};
            const double rx = s.f[SX_RX][i], ry = s.f[SX_RY][i], rz = s.f[SX_RZ][i];
            double acc[7];
            ecp_particle_sums(i, ap, live, acc);
            for (int c = 0; c < 7; ++c) f[c] = f[c] + acc[c];
            // half_kick and ang_kick with the sums in registers
            vx = vx + 0.5 * dt * (f[0] * im);
            vy = vy + 0.5 * dt * (f[1] * im);
            vz = vz + 0.5 * dt * (f[2] * im);
            lx = lx + 0.5 * dt * f[3];
            ly = ly + 0.5 * dt * f[4];
            lz = lz + 0.5 * dt * f[5];
            s.f[SX_FX][i] = f[0]; s.f[SX_FY][i] = f[1]; s.f[SX_FZ][i] = f[2];
            s.f[SX_TX][i] = f[3]; s.f[SX_TY][i] = f[4]; s.f[SX_TZ][i] = f[5];
            pe_acc[i] = f[6];
            s.f[SX_VX][i] = vx; s.f[SX_VY][i] = vy; s.f[SX_VZ][i] = vz;
            s.f[SX_LX][i] = lx; s.f[SX_LY][i] = ly; s.f[SX_LZ][i] = lz;
            v[1] = kin_rot_of(q, lx, ly, lz, ii);
            v[0] = kin_trans_of(mass, vx, vy, vz);
            if (th.on) v[2] = rx * f[0] + ry * f[1] + rz * f[2];
            if (th.pe_acc) v[3] = f[6];
        } else {
            half_kick(s, i, dt);
            ang_kick(s, i, dt);
            v[1] = kin_rot(s, i);
            v[0] = kin_trans(s, i);
            if (th.on) v[2] = virial_of(s, i);
            if (th.pe_acc) v[3] = th.pe_acc[i];
        }
    }
    double tot[4];
    if (!sx_block_reduce_last<4, BS>(v, part, ticket, tot) || threadIdx.x != 0) return;
    scal[SX_KIN_ENG] = tot[0];
    scal[SX_KIN_ROT] = tot[1];
    if (th.pe_acc) scal[SX_POT_ENG] = tot[3];
    if (th.on) thermo_finish(scal, tot[2], th);
    if (dstep) *dstep = *dstep + 1;   // Langevin counter of graph replays
}

// ---------------------------------------------------------------------------
// Nose-Hoover chain (integrators.f90:222-492)
//
// The reference propagates the 3-link chain inside the per-particle loop and
// re-sums sum_k m_k v_k^2 over all particles at every call (O(N^2)).  Here a
// single lane walks the particles in order carrying that sum incrementally
// (the velocity of particle i only changes by the scaling and kick applied
// in its own turn), while the work-group stages per-particle data through
// LDS and applies the resulting updates in parallel.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double exprel_d(double x)
{
    const double g0 = 1.0 / 1, g1 = 1.0 / 6, g2 = 1.0 / 120, g3 = 1.0 / 5040, g4 = 1.0 / 362880;
    if (fabs(x) > 0.01) return (exp(x) - 1.0) / x;
    double h = (x / 2) * (x / 2);
    double f = g4;
    f = f * h + g3;
    f = f * h + g2;
    f = f * h + g1;
    f = f * h + g0;
    return exp(x / 2) * f;
}

// u4_propagator (integrators.f90:274-301) with the j=1 sum passed in
__device__ __forceinline__ void u4(double t, int dir, double sumv, double ndofT, double T, const double hb[3],
                                   double p[3])
{
    for (int q = 0; q < 3; ++q) {
        int j = dir < 0 ? 3 - q : 1 + q;
        double gj = (j == 1) ? (sumv - ndofT) : ((p[j - 2] * p[j - 2] / hb[j - 2]) - T);
        if (j == 3) {
            p[2] = p[2] + t * gj;
        } else {
            double x = t * p[j] / hb[j];
            double c = exprel_d(-x);
            p[j - 1] = p[j - 1] * exp(-x) + t * gj * c;
        }
    }
}

#define NHT_CHUNK 256
// phase 0: nht_initial (chain, scale, chain, kick, drift); phase 1: nht_final
// (kick, chain, scale, chain).  scal holds HB, eta, p_eta; sum of m v^2 is
// computed by this kernel first.
__global__ void __launch_bounds__(NHT_CHUNK) k_nht(DState s, double dt, int phase, int save, double *scal, double ndof,
                                                  double *part)
{
    __shared__ double sv[3][NHT_CHUNK], sa[3][NHT_CHUNK], sm[NHT_CHUNK], sfac[NHT_CHUNK];
    __shared__ double red[NHT_CHUNK];
    __shared__ double sum_sh;
    int t = threadIdx.x;
    int n = s.n;
    // sum of m |v|^2 over all particles before any of this call's updates
    // (in the final step each particle is kicked in its own turn)
    double acc = 0.0;
    for (int i = t; i < n; i += NHT_CHUNK) {
        double vx = s.f[SX_VX][i], vy = s.f[SX_VY][i], vz = s.f[SX_VZ][i];
        acc += s.f[SX_MASS][i] * (vx * vx + vy * vy + vz * vz);
    }
    red[t] = acc;
    __syncthreads();
    for (int o = NHT_CHUNK / 2; o > 0; o >>= 1) {
        if (t < o) red[t] = red[t] + red[t + o];
        __syncthreads();
    }
    if (t == 0) sum_sh = red[0];
    __syncthreads();
    double T = scal[SX_TARGET_TEMP];
    double hb[3] = {scal[SX_HB0], scal[SX_HB1], scal[SX_HB2]};
    double p[3] = {scal[SX_PETA0], scal[SX_PETA1], scal[SX_PETA2]};
    double eta[3] = {scal[SX_ETA0], scal[SX_ETA1], scal[SX_ETA2]};
    double ndofT = ndof * T;
    double tq = dt / 4.0;
    double sumv = sum_sh;
    double ke = 0.0;
    for (int c0 = 0; c0 < n; c0 += NHT_CHUNK) {
        int i = c0 + t;
        if (i < n) {
            double vx = s.f[SX_VX][i], vy = s.f[SX_VY][i], vz = s.f[SX_VZ][i];
            double im = s.f[SX_INV_MASS][i];
            double ax = 0.5 * dt * (s.f[SX_FX][i] * im), ay = 0.5 * dt * (s.f[SX_FY][i] * im);
            double az = 0.5 * dt * (s.f[SX_FZ][i] * im);
            sv[0][t] = vx; sv[1][t] = vy; sv[2][t] = vz;
            sa[0][t] = ax; sa[1][t] = ay; sa[2][t] = az;
            sm[t] = s.f[SX_MASS][i];
        }
        __syncthreads();
        if (t == 0) {
            int cnt = min(NHT_CHUNK, n - c0);
            for (int q = 0; q < cnt; ++q) {
                double m = sm[q];
                double vx = sv[0][q], vy = sv[1][q], vz = sv[2][q];
                if (phase == 1) {
                    // kick first (nht_final_step :374-376)
                    double old = m * (vx * vx + vy * vy + vz * vz);
                    vx = vx + sa[0][q]; vy = vy + sa[1][q]; vz = vz + sa[2][q];
                    sumv = sumv - old + m * (vx * vx + vy * vy + vz * vz);
                }
                u4(tq, -1, sumv, ndofT, T, hb, p);
                double fac = exp(-0.5 * dt * p[0] / hb[0]);
                double before = m * (vx * vx + vy * vy + vz * vz);
                vx = vx * fac; vy = vy * fac; vz = vz * fac;
                sumv = sumv - before + m * (vx * vx + vy * vy + vz * vz);
                for (int k = 0; k < 3; ++k) eta[k] = eta[k] + 0.5 * dt * p[k] / hb[k];
                u4(tq, +1, sumv, ndofT, T, hb, p);
                if (phase == 0) {
                    double mid = m * (vx * vx + vy * vy + vz * vz);
                    vx = vx + sa[0][q]; vy = vy + sa[1][q]; vz = vz + sa[2][q];
                    sumv = sumv - mid + m * (vx * vx + vy * vy + vz * vz);
                }
                sv[0][q] = vx; sv[1][q] = vy; sv[2][q] = vz;
                sfac[q] = fac;
            }
        }
        __syncthreads();
        if (i < n) {
            double vx = sv[0][t], vy = sv[1][t], vz = sv[2][t];
            s.f[SX_VX][i] = vx; s.f[SX_VY][i] = vy; s.f[SX_VZ][i] = vz;
            if (phase == 0) {
                if (save) save_old(s, i);
                s.f[SX_RX][i] = s.f[SX_RX][i] + dt * vx;
                s.f[SX_RY][i] = s.f[SX_RY][i] + dt * vy;
                s.f[SX_RZ][i] = s.f[SX_RZ][i] + dt * vz;
                wrap_pos(s, i);
            } else {
                ke += 0.5 * sm[t] * (vx * vx + vy * vy + vz * vz);
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        scal[SX_PETA0] = p[0]; scal[SX_PETA1] = p[1]; scal[SX_PETA2] = p[2];
        scal[SX_ETA0] = eta[0]; scal[SX_ETA1] = eta[1]; scal[SX_ETA2] = eta[2];
    }
    if (phase == 1) {
        red[t] = ke;
        __syncthreads();
        for (int o = NHT_CHUNK / 2; o > 0; o >>= 1) {
            if (t < o) red[t] = red[t] + red[t + o];
            __syncthreads();
        }
        if (t == 0) scal[SX_KIN_ENG] = red[0];
    }
}

// ---------------------------------------------------------------------------
// The same chain recurrence evaluated in parallel (DEER: Newton iteration on
// the whole sequence).  The particles are cut into at most NHT_CMAX
// consecutive chunks, one per lane, NHT_BLK chunks per workgroup.  A chunk
// maps the chain state entering it, s = (p_eta(1:3), sum of m v^2), to the
// state leaving it; a walk with forward-mode dual numbers gives that map's
// value and 4x4 Jacobian at a guessed entry state.  The affine models of the
// chunks are composed by a two-level prefix scan (a shuffle scan within each
// workgroup, kept; the workgroup totals, published, scanned at the start of
// the next pass), which yields the next Newton iterate of every entry state;
// the iteration converges quadratically to the states the serial walk
// produces (a few passes at equilibrium), and the particle updates follow.
// The results equal the serial recurrence to rounding (the particle updates
// are the same expressions in the same order; only each chunk's entry state
// is reached through the scan).  Each call is one launch (k_nhtp_solve),
// graph-capturable, whose workgroups meet between passes through flags;
// without convergence one workgroup takes the chunks in windows, one window
// after another (nht_windows).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ double texp(double x) { return exp(x); }

// value and one directional derivative
struct D1 {
    double v, d;
};
__device__ __forceinline__ D1 operator+(const D1 &a, const D1 &b) { return D1{a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ D1 operator-(const D1 &a, const D1 &b) { return D1{a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ D1 operator*(const D1 &a, const D1 &b) { return D1{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ D1 operator/(const D1 &a, const D1 &b)
{
    D1 r;
    r.v = a.v / b.v;
    r.d = (a.d - r.v * b.d) * (1.0 / b.v);
    return r;
}
__device__ __forceinline__ D1 operator+(const D1 &a, double b) { return D1{a.v + b, a.d}; }
__device__ __forceinline__ D1 operator-(const D1 &a, double b) { return D1{a.v - b, a.d}; }
__device__ __forceinline__ D1 operator*(const D1 &a, double b) { return D1{a.v * b, a.d * b}; }
__device__ __forceinline__ D1 operator/(const D1 &a, double b) { return a / D1{b, 0.0}; }
__device__ __forceinline__ D1 operator*(double a, const D1 &b) { return b * a; }
__device__ __forceinline__ D1 operator-(double a, const D1 &b) { return D1{a - b.v, -b.d}; }
__device__ __forceinline__ D1 operator-(const D1 &a) { return D1{0.0 - a.v, 0.0 - a.d}; }
__device__ __forceinline__ double val(const D1 &x) { return x.v; }
__device__ __forceinline__ D1 texp(const D1 &x)
{
    double e = exp(x.v);
    return D1{e, e * x.d};
}

// value and all four directional derivatives (the one-wavefront window walks
// of the final kernel)
struct D4 {
    double v, d[4];
};
__device__ __forceinline__ D4 dd(double c) { return D4{c, {0, 0, 0, 0}}; }
__device__ __forceinline__ D4 operator+(const D4 &a, const D4 &b)
{
    return D4{a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2], a.d[3] + b.d[3]}};
}
__device__ __forceinline__ D4 operator-(const D4 &a, const D4 &b)
{
    return D4{a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2], a.d[3] - b.d[3]}};
}
__device__ __forceinline__ D4 operator*(const D4 &a, const D4 &b)
{
    D4 r;
    r.v = a.v * b.v;
    for (int k = 0; k < 4; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
__device__ __forceinline__ D4 operator/(const D4 &a, const D4 &b)
{
    D4 r;
    r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
    for (int k = 0; k < 4; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
__device__ __forceinline__ D4 operator+(const D4 &a, double b) { return a + dd(b); }
__device__ __forceinline__ D4 operator-(const D4 &a, double b) { return a - dd(b); }
__device__ __forceinline__ D4 operator*(const D4 &a, double b)
{
    return D4{a.v * b, {a.d[0] * b, a.d[1] * b, a.d[2] * b, a.d[3] * b}};
}
__device__ __forceinline__ D4 operator/(const D4 &a, double b) { return a / dd(b); }
__device__ __forceinline__ D4 operator*(double a, const D4 &b) { return b * a; }
__device__ __forceinline__ D4 operator-(double a, const D4 &b) { return dd(a) - b; }
__device__ __forceinline__ D4 operator-(const D4 &a) { return dd(0.0) - a; }
__device__ __forceinline__ double val(const D4 &x) { return x.v; }
__device__ __forceinline__ D4 texp(const D4 &x)
{
    const double e = exp(x.v);
    return D4{e, {e * x.d[0], e * x.d[1], e * x.d[2], e * x.d[3]}};
}

// exprel and u4 of the serial kernel above, for doubles or dual numbers
template <class T>
__device__ __forceinline__ T exprel_t(const T &x)
{
    const double g0 = 1.0 / 1, g1 = 1.0 / 6, g2 = 1.0 / 120, g3 = 1.0 / 5040, g4 = 1.0 / 362880;
    if (fabs(val(x)) > 0.01) return (texp(x) - 1.0) / x;
    T h = (x / 2.0) * (x / 2.0);
    T f = h * g4 + g3;
    f = f * h + g2;
    f = f * h + g1;
    f = f * h + g0;
    return texp(x / 2.0) * f;
}

template <class T>
__device__ __forceinline__ void u4_t(double t, int dir, const T &sumv, double ndofT, double Temp, const double hb[3],
                                     T p[3])
{
    for (int q = 0; q < 3; ++q) {
        int j = dir < 0 ? 3 - q : 1 + q;
        T gj = (j == 1) ? (sumv - ndofT) : ((p[j - 2] * p[j - 2] / hb[j - 2]) - Temp);
        if (j == 3) {
            p[2] = p[2] + gj * t;
        } else {
            T x = p[j] * t / hb[j];
            T c = exprel_t(-x);
            p[j - 1] = p[j - 1] * texp(-x) + gj * t * c;
        }
    }
}

// one particle's turn of nht_initial (phase 0) / nht_final (phase 1): chain
// state (p, sumv) in, out; v: the particle's velocity (updated); eta_add: its
// eta increments
template <class T>
__device__ __forceinline__ void nht_turn(int phase, double dt, double tq, double ndofT, double Temp, const double hb[3],
                                         double m, const double a[3], T p[3], T &sumv, T v[3], T eta_add[3])
{
    if (phase == 1) {
        T old = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * m;
        v[0] = v[0] + a[0]; v[1] = v[1] + a[1]; v[2] = v[2] + a[2];
        sumv = sumv - old + (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * m;
    }
    u4_t(tq, -1, sumv, ndofT, Temp, hb, p);
    T fac = texp(p[0] * (-0.5 * dt) / hb[0]);
    T before = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * m;
    v[0] = v[0] * fac; v[1] = v[1] * fac; v[2] = v[2] * fac;
    sumv = sumv - before + (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * m;
    for (int k = 0; k < 3; ++k) eta_add[k] = p[k] * (0.5 * dt) / hb[k];
    u4_t(tq, +1, sumv, ndofT, Temp, hb, p);
    if (phase == 0) {
        T mid = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * m;
        v[0] = v[0] + a[0]; v[1] = v[1] + a[1]; v[2] = v[2] + a[2];
        sumv = sumv - mid + (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * m;
    }
}

#define NHT_CMAX 8192    // most chunks per call
#define NHT_BLK 64       // chunks per workgroup
#define NHT_WGMAX (NHT_CMAX / NHT_BLK)   // 128 workgroups: two workgroup totals per lane in the scans
#define NHT_DIRS 4       // wavefronts per workgroup: one per derivative direction
#define NHT_MAXIT 24     // passes per call (the first at the flat guess); fewer than 63
#define NHT_TOL 1e-12
#define NHT_FLOOR 1e-10
#define NHT_SPIN (1 << 14)   // polls of one wait for the other workgroups before DERR_NHT_SYNC
struct AffMap {
    double M[16];   // row-major 4x4
    double c[4];
};

// (b after a): x -> Mb (Ma x + ca) + cb.  The models only steer the Newton
// iterate, whose converged entry states the particle updates then start
// from, so their sums use fused multiply-adds.
__device__ __forceinline__ AffMap aff_compose(const AffMap &b, const AffMap &a)
{
    AffMap r;
    for (int i = 0; i < 4; ++i) {
        for (int j = 0; j < 4; ++j) {
            double acc = b.M[i * 4] * a.M[j];
            for (int k = 1; k < 4; ++k) acc = fma(b.M[i * 4 + k], a.M[k * 4 + j], acc);
            r.M[i * 4 + j] = acc;
        }
        double acc = b.c[i];
        for (int k = 0; k < 4; ++k) acc = fma(b.M[i * 4 + k], a.c[k], acc);
        r.c[i] = acc;
    }
    return r;
}

__device__ __forceinline__ void aff_apply(const AffMap &A, const double x[4], double y[4])
{
    for (int r = 0; r < 4; ++r) {
        double v = A.c[r];
        for (int k = 0; k < 4; ++k) v = fma(A.M[r * 4 + k], x[k], v);
        y[r] = v;
    }
}

__device__ __forceinline__ AffMap aff_identity()
{
    AffMap r;
    for (int k = 0; k < 16; ++k) r.M[k] = (k % 5 == 0) ? 1.0 : 0.0;
    for (int k = 0; k < 4; ++k) r.c[k] = 0.0;
    return r;
}

__device__ __forceinline__ AffMap aff_load(const double *A)
{
    AffMap r;
    for (int k = 0; k < 16; ++k) r.M[k] = A[k];
    for (int k = 0; k < 4; ++k) r.c[k] = A[16 + k];
    return r;
}

__device__ __forceinline__ void aff_store(const AffMap &a, double *A)
{
    for (int k = 0; k < 16; ++k) A[k] = a.M[k];
    for (int k = 0; k < 4; ++k) A[16 + k] = a.c[k];
}

// A double moved across lanes by a DPP control (gfx9 encodings: row_shr:n
// 0x110 + n, row_bcast:15 0x142, row_bcast:31 0x143, wave_shr:1 0x138);
// lanes outside ROWMASK, or whose source lies outside the row or wave,
// keep `old`.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_d(double v, double old)
{
    const long long bv = __double_as_longlong(v), bo = __double_as_longlong(old);
    const int lo = __builtin_amdgcn_update_dpp((int)bo, (int)bv, CTRL, ROWMASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(bo >> 32), (int)(bv >> 32), CTRL, ROWMASK, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// lane src's double, src wave-uniform
__device__ __forceinline__ double lane_d(double v, int src)
{
    const long long bv = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)bv, src), hi = __builtin_amdgcn_readlane((int)(bv >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// x after the map from the lane the DPP control names (identity where there
// is none; composing with the identity leaves x exactly as it is)
template <int CTRL, int ROWMASK>
__device__ __forceinline__ AffMap aff_dpp_step(const AffMap &x)
{
    AffMap y;
    for (int k = 0; k < 16; ++k) y.M[k] = dpp_d<CTRL, ROWMASK>(x.M[k], (k % 5 == 0) ? 1.0 : 0.0);
    for (int k = 0; k < 4; ++k) y.c[k] = dpp_d<CTRL, ROWMASK>(x.c[k], 0.0);
    return aff_compose(x, y);
}

// inclusive prefix over the lanes of a wavefront, later lanes after earlier:
// within each row of 16 lanes by shifts of 1, 2, 4 and 8, then rows 1 and 3
// after the last lane of the row before, then rows 2 and 3 after lane 31
__device__ __forceinline__ AffMap aff_wave_scan(AffMap x)
{
    x = aff_dpp_step<0x111, 0xf>(x);
    x = aff_dpp_step<0x112, 0xf>(x);
    x = aff_dpp_step<0x114, 0xf>(x);
    x = aff_dpp_step<0x118, 0xf>(x);
    x = aff_dpp_step<0x142, 0xa>(x);
    x = aff_dpp_step<0x143, 0xc>(x);
    return x;
}

// the previous lane's map (the identity in lane 0)
__device__ __forceinline__ AffMap aff_prev_lane(const AffMap &x)
{
    AffMap r;
    for (int k = 0; k < 16; ++k) r.M[k] = dpp_d<0x138, 0xf>(x.M[k], (k % 5 == 0) ? 1.0 : 0.0);
    for (int k = 0; k < 4; ++k) r.c[k] = dpp_d<0x138, 0xf>(x.c[k], 0.0);
    return r;
}

// device scratch of the parallel evaluation.  A workgroup publishes pass k
// of a call by writing its records and then its flag word for the pass's
// parity: the tag epoch * 64 + k + 1 (32 bits, compared modulo 2^32) above
// the largest scaled change of its chunks' entry states in that pass (an
// FP32 value, which is all the convergence test needs), where epoch counts
// the calls made with this buffer.  Every workgroup reads the epoch when it
// starts, and workgroup 0 advances it once every workgroup has published
// pass 0 (so has started), so flags never need clearing.  A workgroup
// publishes pass k + 2 only after every workgroup has published pass k + 1,
// that is after every workgroup has read its pass-k flag.
struct NhtBuf {
    double *T[2];    // by pass parity: element q of workgroup w's composed model at q * NHT_WGMAX + w
    double *msum;    // per workgroup: sum m v^2 over its particles (pass 0)
    double *part;    // per workgroup partial sums of the final reduction (4)
    double *s0;      // [4..6]: chain state after the last particle
    unsigned long long *flag;    // [2][NHT_WGMAX]: per parity and workgroup, tag and change
    unsigned long long *epoch;   // calls evaluated with this buffer
    unsigned *ticket;            // last-workgroup ticket (resets itself)
    int chunk, nb;               // particles per chunk, workgroups (chunks = nb * NHT_BLK)
};

__device__ __forceinline__ void nht_chunk(int n, int chunk, int c, int &i0, int &i1)
{
    i0 = min(n, c * chunk);
    i1 = min(n, i0 + chunk);
}

struct NhtArgs {
    double dt, ndofT, Temp, tq, hb[3];
};

// the chain masses of the call: the coupling given, else n_dof T tau^2 for
// the first link and T tau^2 for the others (tau = 2)
__device__ __forceinline__ void nht_hb(double hb[3], double T, double ndof, double coupling, int has_coupling)
{
    if (has_coupling) {
        hb[0] = hb[1] = hb[2] = coupling;
    } else {
        hb[0] = hb[1] = hb[2] = T * (2.0 * 2.0);
        hb[0] = ndof * T * (2.0 * 2.0);
    }
}

__device__ __forceinline__ void nht_set_hb(double *scal, double ndof, double coupling, int has_coupling)
{
    double hb[3];
    nht_hb(hb, scal[SX_TARGET_TEMP], ndof, coupling, has_coupling);
    scal[SX_HB0] = hb[0]; scal[SX_HB1] = hb[1]; scal[SX_HB2] = hb[2];
}

// largest of two scaled changes; a NaN counts as infinite, so it never converges
__device__ __forceinline__ double nht_dmax(double x, double y)
{
    return (x == x && y == y) ? fmax(x, y) : INFINITY;
}

__device__ __forceinline__ double nht_wave_dmax(double x)
{
    for (int o = SX_WAVE / 2; o > 0; o >>= 1) x = nht_dmax(x, __shfl_xor(x, o, SX_WAVE));
    return x;
}

// The records workgroups exchange between passes (totals, changes, sums of
// m v^2, flags) are written and read only with agent-scope atomic stores
// and loads, which are coherent across the XCDs' L2 caches by themselves (no
// L2 write-back or invalidation, which would cost microseconds per pass and
// evict the particle data).  A workgroup stores its flag only after the
// stores of its record have completed (s_waitcnt), and a reader issues its
// record loads only after its flag loads have returned ready, so a reader
// that sees a flag sees the record published with it.
__device__ __forceinline__ double nht_ld(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void nht_st(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// workgroup w's model from the element-major totals (a wavefront reading
// workgroups 2l and 2l + 1 per lane reads 1 KB in a row per element)
__device__ __forceinline__ AffMap aff_load_shared(const double *T, int w)
{
    AffMap r;
    for (int k = 0; k < 16; ++k) r.M[k] = nht_ld(T + k * NHT_WGMAX + w);
    for (int k = 0; k < 4; ++k) r.c[k] = nht_ld(T + (16 + k) * NHT_WGMAX + w);
    return r;
}

// Waits until every workgroup has published pass `k` (tag >= target modulo
// 2^32; lane l watches workgroups 2l and 2l + 1, both flags loaded before
// either is tested) and returns in d the largest change they published;
// false after NHT_SPIN polls.
__device__ __forceinline__ bool nht_wait(const NhtBuf &B, int k, unsigned target, int lane, double &d)
{
    const int w0 = 2 * lane, w1 = w0 + 1;
    const unsigned long long *F = B.flag + (k & 1) * NHT_WGMAX;
    const unsigned long long *f0 = F + min(w0, B.nb - 1), *f1 = F + min(w1, B.nb - 1);
    for (int it = 0; it < NHT_SPIN; ++it) {
        const unsigned long long a0 = __hip_atomic_load(f0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long a1 = __hip_atomic_load(f1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const bool ok = (int)((unsigned)(a0 >> 32) - target) >= 0 && (int)((unsigned)(a1 >> 32) - target) >= 0;
        if (__all(ok)) {
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            d = nht_wave_dmax(nht_dmax((double)__uint_as_float((unsigned)a0), (double)__uint_as_float((unsigned)a1)));
            return true;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return false;
}

// workgroup b's pass k: its total model, (pass 0) its sum of m v^2; then,
// once those stores have completed, the flag word with its largest change
__device__ __forceinline__ void nht_publish(const NhtBuf &B, int b, int k, unsigned ep, const double *Tw,
                                            double dv, double msum)
{
    // the parity slot picked by a select: an indexed read of the argument
    // struct would be a memory load waited for before the stores
    double *T = ((k & 1) ? B.T[1] : B.T[0]) + b;
    for (int q = 0; q < 20; ++q) nht_st(T + q * NHT_WGMAX, Tw[q]);
    if (k == 0) nht_st(B.msum + b, msum);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    const unsigned long long word = ((unsigned long long)(ep + (unsigned)(k + 1)) << 32) | __float_as_uint((float)dv);
    __hip_atomic_store(B.flag + (k & 1) * NHT_WGMAX + b, word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// particle i's turn of the chain call from chain state (p, sv), with its
// velocity update, its drift in phase 0 and its kinetic energy in phase 1
__device__ __forceinline__ void nht_particle(const DState &s, int i, int phase, int save, const NhtArgs &a,
                                             double p[3], double &sv, double v4[4])
{
    const double dt = a.dt;
    double m = s.f[SX_MASS][i], im = s.f[SX_INV_MASS][i];
    double acc3[3] = {0.5 * dt * (s.f[SX_FX][i] * im), 0.5 * dt * (s.f[SX_FY][i] * im),
                      0.5 * dt * (s.f[SX_FZ][i] * im)};
    double v[3] = {s.f[SX_VX][i], s.f[SX_VY][i], s.f[SX_VZ][i]};
    double ea[3];
    nht_turn(phase, dt, a.tq, a.ndofT, a.Temp, a.hb, m, acc3, p, sv, v, ea);
    for (int q = 0; q < 3; ++q) v4[q] += ea[q];
    s.f[SX_VX][i] = v[0]; s.f[SX_VY][i] = v[1]; s.f[SX_VZ][i] = v[2];
    if (phase == 0) {
        if (save) save_old(s, i);
        s.f[SX_RX][i] = s.f[SX_RX][i] + dt * v[0];
        s.f[SX_RY][i] = s.f[SX_RY][i] + dt * v[1];
        s.f[SX_RZ][i] = s.f[SX_RZ][i] + dt * v[2];
        wrap_pos(s, i);
    } else {
        v4[3] += 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
}

#define NHT_WIN_IT 16   // Newton iterations of one window before its serial walk

// Without convergence of the passes (a call whose chain swings far within
// it): one wavefront takes the chunks in windows of NHT_BLK, one window
// after another from the exact state entering it (s0 for the first).
// Within a window it runs Newton from the flat guess over the window's
// chunks, one per lane, with dual numbers carrying all four derivatives and
// one shuffle scan per iteration (a window spans a short stretch of the
// chain's evolution, so a few iterations converge); a window that has not
// converged after NHT_WIN_IT iterations is walked serially by one lane.
// Then the window's particles are updated from the converged entry states
// and the state after its last chunk enters the next window.
__device__ void nht_windows(const DState &s, int phase, int save, double *scal, double ndof, int chunk,
                            const NhtArgs &a, const double s0[4])
{
    const int lane = threadIdx.x;
    const int nchunks = (s.n + chunk - 1) / chunk;
    const double T = a.Temp;
    const double scale[4] = {sqrt(fabs(a.hb[0] * T)), sqrt(fabs(a.hb[1] * T)), sqrt(fabs(a.hb[2] * T)),
                             fabs(ndof * T)};
    double e[4] = {s0[0], s0[1], s0[2], s0[3]};
    double v4[4] = {0.0, 0.0, 0.0, 0.0};
    for (int w0 = 0; w0 < nchunks; w0 += NHT_BLK) {
        int i0, i1;
        nht_chunk(s.n, chunk, w0 + lane, i0, i1);
        double g[4] = {e[0], e[1], e[2], e[3]};
        bool ok = false;
        for (int it = 0; it < NHT_WIN_IT && !ok; ++it) {
            D4 p[3], sv;
            for (int q = 0; q < 3; ++q) { p[q] = dd(g[q]); p[q].d[q] = 1.0; }
            sv = dd(g[3]);
            sv.d[3] = 1.0;
            for (int i = i0; i < i1; ++i) {
                double m = s.f[SX_MASS][i], im = s.f[SX_INV_MASS][i];
                double acc3[3] = {0.5 * a.dt * (s.f[SX_FX][i] * im), 0.5 * a.dt * (s.f[SX_FY][i] * im),
                                  0.5 * a.dt * (s.f[SX_FZ][i] * im)};
                D4 v[3] = {dd(s.f[SX_VX][i]), dd(s.f[SX_VY][i]), dd(s.f[SX_VZ][i])};
                D4 ea[3];
                nht_turn(phase, a.dt, a.tq, a.ndofT, a.Temp, a.hb, m, acc3, p, sv, v, ea);
            }
            const D4 *out[4] = {&p[0], &p[1], &p[2], &sv};
            AffMap A;
            for (int r = 0; r < 4; ++r) {
                double cr = out[r]->v;
                for (int q = 0; q < 4; ++q) {
                    A.M[r * 4 + q] = out[r]->d[q];
                    cr -= out[r]->d[q] * g[q];
                }
                A.c[r] = cr;
            }
            const AffMap incl = aff_wave_scan(A);
            const AffMap excl = aff_prev_lane(incl);
            double gn[4];
            aff_apply(excl, e, gn);
            double dv = 0.0;
            for (int q = 0; q < 4; ++q) {
                dv = nht_dmax(dv, fabs(gn[q] - g[q]) / (fabs(gn[q]) + scale[q] + 1e-300));
                g[q] = gn[q];
            }
            ok = nht_wave_dmax(dv) < NHT_TOL;
        }
        const int last = min(nchunks, w0 + NHT_BLK) - 1 - w0;   // lane of the window's last chunk
        double x[4] = {0.0, 0.0, 0.0, 0.0};
        if (ok) {
            double p[3] = {g[0], g[1], g[2]}, sv = g[3];
            for (int i = i0; i < i1; ++i) nht_particle(s, i, phase, save, a, p, sv, v4);
            x[0] = p[0]; x[1] = p[1]; x[2] = p[2]; x[3] = sv;
            for (int q = 0; q < 4; ++q) e[q] = lane_d(x[q], last);
        } else {
            int j0, j1, k0, k1;
            nht_chunk(s.n, chunk, w0, j0, k0);
            nht_chunk(s.n, chunk, w0 + last, k1, j1);
            (void)k0; (void)k1;
            if (lane == 0) {
                double p[3] = {e[0], e[1], e[2]}, sv = e[3];
                for (int i = j0; i < j1; ++i) nht_particle(s, i, phase, save, a, p, sv, v4);
                x[0] = p[0]; x[1] = p[1]; x[2] = p[2]; x[3] = sv;
            }
            for (int q = 0; q < 4; ++q) e[q] = lane_d(x[q], 0);
        }
    }
    for (int o = SX_WAVE / 2; o > 0; o >>= 1)
        for (int q = 0; q < 4; ++q) v4[q] = v4[q] + __shfl_xor(v4[q], o, SX_WAVE);
    if (lane == 0) {
        for (int q = 0; q < 3; ++q) scal[SX_ETA0 + q] = scal[SX_ETA0 + q] + v4[q];
        if (phase == 1) scal[SX_KIN_ENG] = v4[3];
        scal[SX_PETA0] = e[0]; scal[SX_PETA1] = e[1]; scal[SX_PETA2] = e[2];
    }
}

// One chain call in one launch of nb <= NHT_WGMAX workgroups, which the
// launch needs resident together (at most 128 of the 256 CUs).  Pass k:
// - wavefront 0 waits until every workgroup has published pass k - 1;
// - for k >= 2 it decides convergence from the largest changes published in
//   passes k - 1 and k - 2: pass k - 1 moved no entry state by more than the
//   tolerance, or by less than NHT_FLOOR without shrinking to a quarter of
//   the change of pass k - 2 (the iteration converges quadratically until
//   the rounding of the composed models stops it, and where the chain
//   amplifies that rounding the floor can lie above NHT_TOL).  Every
//   workgroup reads the same published values, so all decide alike;
// - otherwise it composes the workgroup totals of pass k - 1 by a shuffle
//   scan (two per lane), applies the prefix to the state entering the call
//   and then the chunk's own prefix (kept in registers): the next Newton
//   iterate of every entry state.  Pass 0 starts from the flat guess (the
//   chain momenta entering the call, sum m v^2 = n_dof T) and sums m v^2 per
//   workgroup, which pass 1 adds in a fixed order into the state entering
//   the call;
// - a workgroup none of whose entry states moved by more than the tolerance
//   keeps its models; the others walk their chunks with dual numbers, four
//   wavefronts each carrying the derivative along one component of the entry
//   state, which gives every chunk's affine model x -> J x + (E - J g), and a
//   shuffle scan over the workgroup's chunks gives each chunk's prefix and
//   the workgroup total, which is published.
// Converged: the particle updates of every chunk from its entry state with
// the serial recurrence's expressions in its order, then the sums of the eta
// increments and kinetic energy (fixed order) and the chain state after the
// last particle.  Not converged after `maxit` passes: workgroup 0 takes the
// chunks in windows (nht_windows).  A wait that outlasts NHT_SPIN polls (a
// workgroup that never became resident) raises DERR_NHT_SYNC and ends the
// workgroup; every wait is bounded, so the launch always ends.
__global__ void __launch_bounds__(NHT_BLK * NHT_DIRS) k_nhtp_solve(DState s, double dt, int phase, int save,
                                                                   double *scal, double ndof, NhtBuf B, int maxit,
                                                                   unsigned long long *stats, double coupling,
                                                                   int has_coupling, int *err)
{
    __shared__ double gsh[NHT_BLK][4], col[NHT_DIRS][NHT_BLK][4], val[NHT_BLK][4];
    __shared__ double psh[NHT_BLK][20], tsh[20];   // each chunk's prefix within the workgroup, the total
    __shared__ int ctl_sh[2];
    const int b = blockIdx.x, lane = threadIdx.x % NHT_BLK, dir = threadIdx.x / NHT_BLK;
    const int c = b * NHT_BLK + lane;
    NhtArgs a;
    a.dt = dt;
    a.Temp = scal[SX_TARGET_TEMP];
    a.ndofT = ndof * a.Temp;
    a.tq = dt / 4.0;
    nht_hb(a.hb, a.Temp, ndof, coupling, has_coupling);
    int i0, i1;
    nht_chunk(s.n, B.chunk, c, i0, i1);
    const unsigned long long epoch = B.epoch[0];
    const unsigned ep = (unsigned)(epoch * 64ull);
    // wavefront 0: the chunk's entry state, the state entering the call
    double g[4] = {scal[SX_PETA0], scal[SX_PETA1], scal[SX_PETA2], a.ndofT};
    double s0[4] = {g[0], g[1], g[2], 0.0};
    double dprev = INFINITY, msum = 0.0, dvk = 0.0;
    const double scale[4] = {sqrt(fabs(a.hb[0] * a.Temp)), sqrt(fabs(a.hb[1] * a.Temp)),
                             sqrt(fabs(a.hb[2] * a.Temp)), fabs(ndof * a.Temp)};
    int ctl;   // 0 walk, 1 keep the models, 2 converged, 3 not converged, 4 waited in vain
    for (int k = 0;; ++k) {
        if (dir == 0) {
            ctl = 0;
            dvk = 0.0;
            double d = INFINITY;
            if (k == 0) {
                for (int i = i0; i < i1; ++i) {
                    const double vx = s.f[SX_VX][i], vy = s.f[SX_VY][i], vz = s.f[SX_VZ][i];
                    msum += s.f[SX_MASS][i] * (vx * vx + vy * vy + vz * vz);
                }
                for (int o = SX_WAVE / 2; o > 0; o >>= 1) msum = msum + __shfl_xor(msum, o, SX_WAVE);
                if (b == 0 && lane == 0 && stats) atomicAdd(&stats[1], 1ull);
            } else if (!nht_wait(B, k - 1, ep + (unsigned)k, lane, d)) {
                if (lane == 0) atomicCAS(err, DERR_NONE, DERR_NHT_SYNC);
                ctl = 4;
            } else {
                if (k == 1) d = INFINITY;   // pass 0 has no previous iterate
                // converged (see above), or the observed contraction makes the
                // next step smaller than the tolerance: d * (d / dprev), the
                // estimate of a linearly converging iteration, which bounds a
                // quadratically converging one.  Only once both changes are
                // measured (k >= 3) and the previous one was already below the
                // thermal scale: after a diverging pass (dprev infinite) the
                // estimate says nothing.
                const bool conv = k >= 2 && (d < NHT_TOL || (d < NHT_FLOOR && d > 0.25 * dprev) ||
                                             (k >= 3 && dprev < 1.0 && d < 1e-3 * dprev && d * (d / dprev) < NHT_TOL));
                dprev = d;
                if (conv) ctl = 2;
                else if (k == maxit) ctl = 3;
            }
            // the models of pass k - 1 for the next iterate; in pass 1 also the
            // sums of m v^2 (needed by the windows too, when one pass is all)
            if (k >= 1 && (ctl == 0 || (k == 1 && ctl == 3))) {
                const int w0 = 2 * lane, w1 = w0 + 1;
                const double *Tp = ((k - 1) & 1) ? B.T[1] : B.T[0];
                const bool ld = ctl == 0;
                const AffMap t0 = ld && w0 < B.nb ? aff_load_shared(Tp, w0) : aff_identity();
                const AffMap t1 = ld && w1 < B.nb ? aff_load_shared(Tp, w1) : aff_identity();
                if (k == 1) {
                    double m2 = (w0 < B.nb ? nht_ld(B.msum + w0) : 0.0) + (w1 < B.nb ? nht_ld(B.msum + w1) : 0.0);
                    for (int o = SX_WAVE / 2; o > 0; o >>= 1) m2 = m2 + __shfl_xor(m2, o, SX_WAVE);
                    s0[3] = m2;
                    if (b == 0 && lane == 0) {
                        scal[SX_HB0] = a.hb[0]; scal[SX_HB1] = a.hb[1]; scal[SX_HB2] = a.hb[2];
                        B.epoch[0] = epoch + 1ull;   // every workgroup has started
                    }
                }
                if (ctl == 0) {
                    const AffMap x = aff_wave_scan(aff_compose(t1, t0));
                    double e[4], eprev[4], f[4], E[4], gn[4];
                    aff_apply(x, s0, e);   // state after workgroups 0 .. 2 lane + 1
                    for (int q = 0; q < 4; ++q) eprev[q] = dpp_d<0x138, 0xf>(e[q], s0[q]);
                    aff_apply(t0, eprev, f);   // state after workgroups 0 .. 2 lane
                    for (int q = 0; q < 4; ++q) E[q] = lane_d((b & 1) ? f[q] : eprev[q], b >> 1);
                    aff_apply(aff_load(psh[lane]), E, gn);
                    double dv = 0.0;
                    for (int q = 0; q < 4; ++q) {
                        dv = nht_dmax(dv, fabs(gn[q] - g[q]) / (fabs(gn[q]) + scale[q] + 1e-300));
                        g[q] = gn[q];
                    }
                    dvk = nht_wave_dmax(dv);
                    ctl = dvk < NHT_TOL ? 1 : 0;
                    if (b == 0 && lane == 0 && stats) atomicAdd(&stats[0], 1ull);
                }
            }
            for (int q = 0; q < 4; ++q) gsh[lane][q] = g[q];
            if (lane == 0) ctl_sh[k & 1] = ctl;
        }
        __syncthreads();
        ctl = ctl_sh[k & 1];
        if (ctl >= 2) break;
        if (ctl == 0) {
            // the walk: wavefront dir carries the derivative along entry component dir
            D1 p[3], sv;
            for (int q = 0; q < 3; ++q) p[q] = D1{gsh[lane][q], q == dir ? 1.0 : 0.0};
            sv = D1{gsh[lane][3], dir == 3 ? 1.0 : 0.0};
            for (int i = i0; i < i1; ++i) {
                double m = s.f[SX_MASS][i], im = s.f[SX_INV_MASS][i];
                double acc3[3] = {0.5 * dt * (s.f[SX_FX][i] * im), 0.5 * dt * (s.f[SX_FY][i] * im),
                                  0.5 * dt * (s.f[SX_FZ][i] * im)};
                D1 v[3] = {D1{s.f[SX_VX][i], 0.0}, D1{s.f[SX_VY][i], 0.0}, D1{s.f[SX_VZ][i], 0.0}};
                D1 ea[3];
                nht_turn(phase, dt, a.tq, a.ndofT, a.Temp, a.hb, m, acc3, p, sv, v, ea);
            }
            const D1 out[4] = {p[0], p[1], p[2], sv};
            for (int r = 0; r < 4; ++r) {
                col[dir][lane][r] = out[r].d;
                if (dir == 0) val[lane][r] = out[r].v;
            }
            __syncthreads();
            if (dir == 0) {
                AffMap A;
                for (int r = 0; r < 4; ++r) {
                    double cr = val[lane][r];
                    for (int q = 0; q < 4; ++q) {
                        A.M[r * 4 + q] = col[q][lane][r];
                        cr -= col[q][lane][r] * g[q];
                    }
                    A.c[r] = cr;
                }
                const AffMap incl = aff_wave_scan(A);
                aff_store(aff_prev_lane(incl), psh[lane]);
                if (lane == NHT_BLK - 1) aff_store(incl, tsh);
            }
        }
        // published by the lane that stored the total (its own LDS writes)
        if (dir == 0 && lane == NHT_BLK - 1) nht_publish(B, b, k, ep, tsh, dvk, msum);
    }
    if (ctl == 4) return;
    if (ctl == 3) {
        if (b == 0 && dir == 0) {
            if (lane == 0 && stats) atomicAdd(&stats[2], 1ull);
            nht_windows(s, phase, save, scal, ndof, B.chunk, a, s0);
        }
        return;
    }
    if (dir != 0) return;
    double v4[4] = {0.0, 0.0, 0.0, 0.0};   // eta increments, kinetic energy
    double p[3] = {g[0], g[1], g[2]}, sv = g[3];
    for (int i = i0; i < i1; ++i) nht_particle(s, i, phase, save, a, p, sv, v4);
    // the chain state after the last particle belongs to the last non-empty chunk
    if (i1 == s.n && i0 < s.n)
        for (int q = 0; q < 3; ++q) nht_st(B.s0 + 4 + q, p[q]);
    for (int o = SX_WAVE / 2; o > 0; o >>= 1)
        for (int q = 0; q < 4; ++q) v4[q] = v4[q] + __shfl_xor(v4[q], o, SX_WAVE);
    if (lane == 0)
        for (int q = 0; q < 4; ++q) nht_st(B.part + b * 4 + q, v4[q]);
    // the sums by the last workgroup to arrive (same protocol as the passes:
    // the stores complete before the ticket, the loads follow it)
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    unsigned t = 0;
    if (lane == 0) t = __hip_atomic_fetch_add(B.ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    t = __shfl(t, 0, SX_WAVE);
    if (t != (unsigned)(B.nb - 1)) return;
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    const int w0 = 2 * lane, w1 = w0 + 1;
    double tot[4];
    for (int q = 0; q < 4; ++q) {
        tot[q] = (w0 < B.nb ? nht_ld(B.part + w0 * 4 + q) : 0.0) + (w1 < B.nb ? nht_ld(B.part + w1 * 4 + q) : 0.0);
        for (int o = SX_WAVE / 2; o > 0; o >>= 1) tot[q] = tot[q] + __shfl_xor(tot[q], o, SX_WAVE);
    }
    if (lane == 0) {
        for (int q = 0; q < 3; ++q) scal[SX_ETA0 + q] = scal[SX_ETA0 + q] + tot[q];
        if (phase == 1) scal[SX_KIN_ENG] = tot[3];
        for (int q = 0; q < 3; ++q) scal[SX_PETA0 + q] = nht_ld(B.s0 + 4 + q);
        __hip_atomic_store(B.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ void k_nht_hb(double *scal, double ndof, double coupling, int has_coupling)
{
    if (threadIdx.x != 0) return;
    nht_set_hb(scal, ndof, coupling, has_coupling);
}

// ---------------------------------------------------------------------------
// RATTLE: one molecule per wavefront, bonds swept in the reference's order
// ---------------------------------------------------------------------------
__device__ __forceinline__ double shfl_d(double v, int src) { return __shfl(v, src, SX_WAVE); }

template <int PHASE>
__global__ void __launch_bounds__(64) k_rattle(DState s, double dt, int nmol, const int *atom_start, const int *atoms,
                                              const int *bond_start, const int *bond_a, const int *bond_b,
                                              const double *bond_len, double *vir_part, int *err)
{
    int m = blockIdx.x;
    int lane = threadIdx.x;
    double vir = 0.0;
    if (m < nmol) {
        int a0 = atom_start[m], na = atom_start[m + 1] - a0;
        int b0 = bond_start[m], nb = bond_start[m + 1] - b0;
        // lane k holds atom k of the molecule (na <= 64)
        int g = lane < na ? atoms[a0 + lane] : 0;
        double rx = 0, ry = 0, rz = 0, vx = 0, vy = 0, vz = 0, ox = 0, oy = 0, oz = 0, im = 0;
        if (lane < na) {
            rx = s.f[SX_RX][g]; ry = s.f[SX_RY][g]; rz = s.f[SX_RZ][g];
            vx = s.f[SX_VX][g]; vy = s.f[SX_VY][g]; vz = s.f[SX_VZ][g];
            ox = s.f[SX_RX_OLD][g]; oy = s.f[SX_RY_OLD][g]; oz = s.f[SX_RZ_OLD][g];
            im = s.f[SX_INV_MASS][g];
        }
        bool moved = true;
        const double r_tol = 1.0e-5;
        int failed = 0;
        for (int rattle = 0; rattle < 100; ++rattle) {
            if (!__any(lane < na && moved)) break;
            for (int q = 0; q < nb; ++q) {
                int la = bond_a[b0 + q], lb = bond_b[b0 + q];
                double len = bond_len[b0 + q];
                bool ma = __shfl((int)moved, la, SX_WAVE) != 0, mb = __shfl((int)moved, lb, SX_WAVE) != 0;
                if (!(ma || mb)) continue;
                double ax = shfl_d(rx, la), ay = shfl_d(ry, la), az = shfl_d(rz, la);
                double bx = shfl_d(rx, lb), by = shfl_d(ry, lb), bz = shfl_d(rz, lb);
                double R[3] = {ax - bx, ay - by, az - bz};
                if (s.periodic)
                    for (int k = 0; k < 3; ++k) R[k] = sx_mimg(R[k], s.box[k]);
                double ima = shfl_d(im, la), imb = shfl_d(im, lb);
                if (PHASE == 0) {
                    double chi = len * len - (R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
                    if (fabs(chi) > 2 * r_tol * (len * len)) {
                        double Ro[3] = {shfl_d(ox, la) - shfl_d(ox, lb), shfl_d(oy, la) - shfl_d(oy, lb),
                                        shfl_d(oz, la) - shfl_d(oz, lb)};
                        if (s.periodic)
                            for (int k = 0; k < 3; ++k) Ro[k] = sx_mimg(Ro[k], s.box[k]);
                        double vdot = R[0] * Ro[0] + R[1] * Ro[1] + R[2] * Ro[2];
                        if (vdot < r_tol * (len * len)) { failed = 1; break; }
                        double L = chi / (2.0 * vdot * (ima + imb));
                        vir = vir + (L * (len * len)) / (dt * dt);
                        if (lane == la) {
                            double c = L * ima;
                            rx = rx + c * Ro[0]; ry = ry + c * Ro[1]; rz = rz + c * Ro[2];
                            vx = vx + c * Ro[0] / dt; vy = vy + c * Ro[1] / dt; vz = vz + c * Ro[2] / dt;
                        }
                        if (lane == lb) {
                            double c = L * imb;
                            rx = rx - c * Ro[0]; ry = ry - c * Ro[1]; rz = rz - c * Ro[2];
                            vx = vx - c * Ro[0] / dt; vy = vy - c * Ro[1] / dt; vz = vz - c * Ro[2] / dt;
                        }
                        if (lane == la || lane == lb) moved = true;
                    } else {
                        if (lane == la || lane == lb) moved = false;
                    }
                } else {
                    double V[3] = {shfl_d(vx, la) - shfl_d(vx, lb), shfl_d(vy, la) - shfl_d(vy, lb),
                                   shfl_d(vz, la) - shfl_d(vz, lb)};
                    double vdot = R[0] * V[0] + R[1] * V[1] + R[2] * V[2];
                    double L = -vdot / ((ima + imb) * (len * len));
                    if (fabs(L) > r_tol) {
                        vir = vir + (L * (len * len)) / dt;
                        if (lane == la) {
                            double c = L * ima;
                            vx = vx + c * R[0]; vy = vy + c * R[1]; vz = vz + c * R[2];
                        }
                        if (lane == lb) {
                            double c = L * imb;
                            vx = vx - c * R[0]; vy = vy - c * R[1]; vz = vz - c * R[2];
                        }
                        if (lane == la || lane == lb) moved = true;
                    } else {
                        if (lane == la || lane == lb) moved = false;
                    }
                }
            }
            if (failed) break;
        }
        if (failed && lane == 0) atomicCAS(err, 0, DERR_CONSTRAINT);
        if (lane < na) {
            s.f[SX_RX][g] = rx; s.f[SX_RY][g] = ry; s.f[SX_RZ][g] = rz;
            s.f[SX_VX][g] = vx; s.f[SX_VY][g] = vy; s.f[SX_VZ][g] = vz;
        }
    }
    if (lane == 0) vir_part[blockIdx.x] = vir;
}

// RATTLE for molecules of at most 64 bonds with the sweeps pipelined.  Bond
// q of sweep s runs at step lvl[q] + P*s (the schedule sx_set_constraints
// derives from the bond order): the bonds of one step touch disjoint atoms,
// and every atom sees the updates that touch it in the reference's serial
// order at strictly increasing steps, so positions, velocities and moved
// flags take exactly the values of the serial sweeps.  A chain whose bonds
// are listed along it has lvl[q] = q and P = 2: S sweeps take nb + 2(S-1)
// steps instead of nb*S dependent bond updates.
//
// The reference runs another sweep while some atom is flagged.  A sweep
// without corrections leaves every flag clear (an atom flagged at its start
// is evaluated, and cleared, by the first bond of the sweep that touches
// it), and a sweep that starts with every flag clear skips every bond.  So
// the wavefront stops at the end of the first sweep without corrections:
// every sweep the reference runs is run, and any sweep run beyond them (and
// any step already run for a later sweep) starts from clear flags and
// changes nothing.  Lane q records in bit s mod 64 of a mask whether bond q
// was corrected in sweep s; at the step that completes sweep s the wavefront
// stops if no lane's bit is set.  At most span/P + 1 <= 64 sweeps are in
// flight, so the bits do not collide.  Atom state lives in LDS, one
// wavefront per workgroup.
template <int PHASE>
__global__ void __launch_bounds__(64) k_rattle_pipe(DState s, double dt, const int *atom_start, const int *atoms,
                                                   const int *bond_start, const int *bond_a, const int *bond_b,
                                                   const double *bond_len, const int *bond_lvl, const int *sched,
                                                   double *vir_part, int *err)
{
    __shared__ double sr[3][SX_WAVE], sv[3][SX_WAVE];
    __shared__ int smv[SX_WAVE];
    const int m = blockIdx.x, lane = threadIdx.x;
    const int a0 = atom_start[m], na = atom_start[m + 1] - a0;
    const int b0 = bond_start[m], nb = bond_start[m + 1] - b0;
    const int g = lane < na ? atoms[a0 + lane] : 0;
    double rx = 0, ry = 0, rz = 0, ox = 0, oy = 0, oz = 0, im = 0;
    if (lane < na) {
        rx = s.f[SX_RX][g]; ry = s.f[SX_RY][g]; rz = s.f[SX_RZ][g];
        sr[0][lane] = rx; sr[1][lane] = ry; sr[2][lane] = rz;
        sv[0][lane] = s.f[SX_VX][g]; sv[1][lane] = s.f[SX_VY][g]; sv[2][lane] = s.f[SX_VZ][g];
        if (PHASE == 0) { ox = s.f[SX_RX_OLD][g]; oy = s.f[SX_RY_OLD][g]; oz = s.f[SX_RZ_OLD][g]; }
        im = s.f[SX_INV_MASS][g];
    }
    smv[lane] = 1;   // moved = .TRUE.
    int qa = 0, qb = 0, next = INT_MAX;
    double qlen = 0.0;
    if (lane < nb) {
        qa = bond_a[b0 + lane]; qb = bond_b[b0 + lane]; qlen = bond_len[b0 + lane];
        next = bond_lvl[b0 + lane];
    }
    const double ima = shfl_d(im, qa), imb = shfl_d(im, qb), len = qlen;
    // the bond vector fixed during the phase: old positions in a, positions in b
    double Q[3] = {PHASE == 0 ? shfl_d(ox, qa) - shfl_d(ox, qb) : shfl_d(rx, qa) - shfl_d(rx, qb),
                   PHASE == 0 ? shfl_d(oy, qa) - shfl_d(oy, qb) : shfl_d(ry, qa) - shfl_d(ry, qb),
                   PHASE == 0 ? shfl_d(oz, qa) - shfl_d(oz, qb) : shfl_d(rz, qa) - shfl_d(rz, qb)};
    if (s.periodic)
        for (int k = 0; k < 3; ++k) Q[k] = sx_mimg(Q[k], s.box[k]);
    const int P = sched[2 * m], span = sched[2 * m + 1];
    const double r_tol = 1.0e-5;
    double vir = 0.0;
    unsigned long long corr = 0ull;
    int sw = 0, failed = 0;
    int chk = span, S = 0;   // the step that completes sweep S
    __syncthreads();
    for (int t = 0;; ++t) {
        if (t == next) {
            const int ma = smv[qa], mb = smv[qb];
            const double va[3] = {sv[0][qa], sv[1][qa], sv[2][qa]}, vb[3] = {sv[0][qb], sv[1][qb], sv[2][qb]};
            if (PHASE == 0) {
                const double ra[3] = {sr[0][qa], sr[1][qa], sr[2][qa]}, rb[3] = {sr[0][qb], sr[1][qb], sr[2][qb]};
                if (ma | mb) {
                    double R[3] = {ra[0] - rb[0], ra[1] - rb[1], ra[2] - rb[2]};
                    if (s.periodic)
                        for (int k = 0; k < 3; ++k) R[k] = sx_mimg(R[k], s.box[k]);
                    const double chi = len * len - (R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
                    if (fabs(chi) > 2 * r_tol * (len * len)) {
                        const double vdot = R[0] * Q[0] + R[1] * Q[1] + R[2] * Q[2];
                        if (vdot < r_tol * (len * len)) {
                            failed = 1;
                        } else {
                            const double L = chi / (2.0 * vdot * (ima + imb));
                            vir = vir + (L * (len * len)) / (dt * dt);
                            // atom b's update as the addition of the negated
                            // product (negation is exact)
                            const double ca = L * ima, cb = -(L * imb);
                            for (int k = 0; k < 3; ++k) {
                                sr[k][qa] = ra[k] + ca * Q[k];
                                sr[k][qb] = rb[k] + cb * Q[k];
                                sv[k][qa] = va[k] + ca * Q[k] / dt;
                                sv[k][qb] = vb[k] + cb * Q[k] / dt;
                            }
                            smv[qa] = 1; smv[qb] = 1;
                            corr |= 1ull << (sw & 63);
                        }
                    } else {
                        smv[qa] = 0; smv[qb] = 0;
                    }
                }
            } else if (ma | mb) {
                const double V[3] = {va[0] - vb[0], va[1] - vb[1], va[2] - vb[2]};
                const double vdot = Q[0] * V[0] + Q[1] * V[1] + Q[2] * V[2];
                const double L = -vdot / ((ima + imb) * (len * len));
                if (fabs(L) > r_tol) {
                    vir = vir + (L * (len * len)) / dt;
                    const double ca = L * ima, cb = -(L * imb);
                    for (int k = 0; k < 3; ++k) {
                        sv[k][qa] = va[k] + ca * Q[k];
                        sv[k][qb] = vb[k] + cb * Q[k];
                    }
                    smv[qa] = 1; smv[qb] = 1;
                    corr |= 1ull << (sw & 63);
                } else {
                    smv[qa] = 0; smv[qb] = 0;
                }
            }
            // the reference's cap: sweeps 0..99
            next = ++sw < 100 ? next + P : INT_MAX;
        }
        __syncthreads();
        if (__any(failed)) break;
        if (t == chk) {
            const bool more = __any((corr >> (S & 63)) & 1ull);
            corr &= ~(1ull << (S & 63));
            if (!more || S == 99) break;
            ++S;
            chk += P;
        }
    }
    if (failed) atomicCAS(err, 0, DERR_CONSTRAINT);
    if (lane < na) {
        s.f[SX_RX][g] = sr[0][lane]; s.f[SX_RY][g] = sr[1][lane]; s.f[SX_RZ][g] = sr[2][lane];
        s.f[SX_VX][g] = sv[0][lane]; s.f[SX_VY][g] = sv[1][lane]; s.f[SX_VZ][g] = sv[2][lane];
    }
    for (int o = SX_WAVE / 2; o > 0; o >>= 1) vir = vir + __shfl_xor(vir, o, SX_WAVE);
    if (lane == 0) vir_part[m] = vir;
}

// ---------------------------------------------------------------------------
// bonded: per-term evaluation into a buffer, per-atom gather
// ---------------------------------------------------------------------------
__device__ __forceinline__ void bvec(const DState &s, int a, int b, double v[3])
{
    v[0] = s.f[SX_RX][b] - s.f[SX_RX][a];
    v[1] = s.f[SX_RY][b] - s.f[SX_RY][a];
    v[2] = s.f[SX_RZ][b] - s.f[SX_RZ][a];
    if (s.periodic)
        for (int k = 0; k < 3; ++k) v[k] = sx_mimg(v[k], s.box[k]);
}

// bends: kind 1 bend_ (force.f90:2244), 2 cosine_bend_ (:2345), 5 angle_bend_
// (:2464, harmonic in theta); dihedrals: 3 dih_ (:2286), 4 cosine_dih_
// (:2392), 6 angle_dih_ (:2508).  th0 = pi - angle in radians, cos0/sin0 its
// cosine and sine (computed once on the host).
__device__ __forceinline__ bool bonded_is_bend(int kind) { return kind <= 2 || kind == 5; }

__global__ void k_bonded_terms(DState s, int kind, int count, const int *idx, double coeff, double th0,
                               double cos0, double sin0, double *out)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int *q = idx + 4 * t;
    double *o = out + 13L * t;
    if (bonded_is_bend(kind)) {
        double VI[3], VJ[3];
        bvec(s, q[0], q[1], VI);
        bvec(s, q[1], q[2], VJ);
        double CC11 = VI[0] * VI[0] + VI[1] * VI[1] + VI[2] * VI[2];
        double CC12 = VI[0] * VJ[0] + VI[1] * VJ[1] + VI[2] * VJ[2];
        double CC22 = VJ[0] * VJ[0] + VJ[1] * VJ[1] + VJ[2] * VJ[2];
        double prefac = 1.0 / sqrt(CC11 * CC22);
        double fac = CC12, fac1 = fac / CC22, fac2 = fac / CC11;
        double strength, pe;
        if (kind == 1) {
            pe = coeff * (1.0 - prefac * fac);
            strength = coeff;
        } else if (kind == 5) {
            double theta = acos(prefac * fac);
            strength = coeff * (theta - th0) / sin(theta);
            pe = (coeff / 2.0) * (theta - th0) * (theta - th0);
        } else {
            double cos_theta = prefac * fac;
            strength = coeff * (cos_theta - cos0);
            pe = (coeff / 2.0) * (cos_theta - cos0) * (cos_theta - cos0);
        }
        o[0] = pe;
        for (int k = 0; k < 3; ++k) {
            o[1 + k] = strength * prefac * (fac2 * VI[k] - VJ[k]);
            o[4 + k] = strength * prefac * (fac1 * VJ[k] - fac2 * VI[k] + VJ[k] - VI[k]);
            o[7 + k] = -strength * prefac * (fac1 * VJ[k] - VI[k]);
            o[10 + k] = 0.0;
        }
    } else {
        double VI[3], VJ[3], VK[3];
        bvec(s, q[0], q[1], VI);
        bvec(s, q[1], q[2], VJ);
        bvec(s, q[2], q[3], VK);
        double CC11 = VI[0] * VI[0] + VI[1] * VI[1] + VI[2] * VI[2];
        double CC12 = VI[0] * VJ[0] + VI[1] * VJ[1] + VI[2] * VJ[2];
        double CC13 = VI[0] * VK[0] + VI[1] * VK[1] + VI[2] * VK[2];
        double CC22 = VJ[0] * VJ[0] + VJ[1] * VJ[1] + VJ[2] * VJ[2];
        double CC23 = VJ[0] * VK[0] + VJ[1] * VK[1] + VJ[2] * VK[2];
        double CC33 = VK[0] * VK[0] + VK[1] * VK[1] + VK[2] * VK[2];
        double DD12 = CC11 * CC22 - CC12 * CC12;
        double DD23 = CC22 * CC33 - CC23 * CC23;
        double prefac = 1.0 / sqrt(DD23 * DD12);
        double fac = CC23 * CC12 - CC13 * CC22;
        double fac1 = fac / DD23, fac2 = fac / DD12;
        double strength, pe;
        if (kind == 3) {
            pe = coeff * (1.0 + prefac * fac);
            strength = coeff;
        } else {
            double vixvj[3];
            cross3(VI, VJ, vixvj);
            double fac_sin = vixvj[0] * VK[0] + vixvj[1] * VK[1] + vixvj[2] * VK[2];
            double cos_phi = -prefac * fac;
            double sin_phi = prefac * fac_sin * sqrt(VJ[0] * VJ[0] + VJ[1] * VJ[1] + VJ[2] * VJ[2]);
            if (kind == 6) {
                strength = coeff * (cos0 - sin0 * (cos_phi / sin_phi));
                pe = coeff * (1.0 - cos_phi * cos0 - sin_phi * sin0);
            } else {
                double sg = atan2(sin_phi, cos_phi);
                strength = coeff * (cos(sg) - cos0);
                pe = (coeff / 2.0) * (cos(sg) - cos0) * (cos(sg) - cos0);
            }
        }
        o[0] = pe;
        for (int k = 0; k < 3; ++k) {
            o[10 + k] = -strength * prefac * (CC12 * VJ[k] - CC22 * VI[k] - fac1 * (CC22 * VK[k] - CC23 * VJ[k]));
            o[7 + k] = -strength * prefac * (CC12 * VK[k] - CC12 * VJ[k] + CC23 * VI[k] + CC22 * VI[k] - 2.0 * CC13 * VJ[k]
                                            - fac2 * (CC11 * VJ[k] - CC12 * VI[k]) - fac1 * (CC33 * VJ[k] - CC22 * VK[k] - CC23 * VK[k] + CC23 * VJ[k]));
            o[4 + k] = -strength * prefac * (-CC12 * VK[k] + CC23 * VJ[k] - CC23 * VI[k] - CC22 * VK[k] + 2.0 * CC13 * VJ[k]
                                            - fac2 * (CC22 * VI[k] - CC11 * VJ[k] - CC12 * VJ[k] + CC12 * VI[k]) - fac1 * (-CC33 * VJ[k] + CC23 * VK[k]));
            o[1 + k] = -strength * prefac * (-CC23 * VJ[k] + CC22 * VK[k] - fac2 * (-CC22 * VI[k] + CC12 * VJ[k]));
        }
    }
}

// atom_term entries: (global term slot * 4 + position), slots index a
// concatenation of all registered term buffers
__global__ void __launch_bounds__(256) k_bonded_gather(DState s, const int *atom_start, const int *atom_term,
                                                       const double *const *bufs, const long *buf_off, int nbufs,
                                                       double *pe_acc)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (i < s.n) {
        double fx = 0, fy = 0, fz = 0;
        for (int q = atom_start[i]; q < atom_start[i + 1]; ++q) {
            int code = atom_term[q];
            long slot = code >> 2;
            int pos = code & 3;
            int b = 0;
            while (b + 1 < nbufs && slot >= buf_off[b + 1]) ++b;
            const double *o = bufs[b] + 13L * (slot - buf_off[b]);
            fx = fx + o[1 + 3 * pos]; fy = fy + o[2 + 3 * pos]; fz = fz + o[3 + 3 * pos];
            if (pos == 0) v[0] = v[0] + o[0];
        }
        s.f[SX_FX][i] = s.f[SX_FX][i] + fx;
        s.f[SX_FY][i] = s.f[SX_FY][i] + fy;
        s.f[SX_FZ][i] = s.f[SX_FZ][i] + fz;
        pe_acc[i] = pe_acc[i] + v[0];
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static void reduce_into(sx_ctx *c, int nb, int width, std::initializer_list<int> ids, int accumulate)
{
    std::vector<int> v(ids);
    sx_reduce_to_scalar(c, nb, width, v.data(), (int)v.size(), accumulate);
}

void sx_vv_initial_impl(sx_ctx *c, double dt, bool save)
{
    hipLaunchKernelGGL(k_vv_initial, dim3(sx_blocks(c->n, 256)), dim3(256), 0, c->stream, c->dstate(), dt, save ? 1 : 0);
}

static ThermoArgs thermo_args(sx_ctx *c, bool on)
{
    ThermoArgs th;
    th.on = on ? 1 : 0;
    th.nc = c->nc;
    th.nrigid = c->nrigid;
    th.volume = c->box[0] * c->box[1] * c->box[2];
    th.pe_acc = (on && c->defer_pe) ? c->d_pe_acc : nullptr;
    return th;
}

void sx_vv_final_impl(sx_ctx *c, double dt, bool thermo)
{
    int nb = sx_blocks(c->n, 256);
    sx_ensure_part(c, nb, 3);
    hipLaunchKernelGGL(k_vv_final, dim3(nb), dim3(256), 0, c->stream, c->dstate(), dt, c->d_part, c->d_ticket,
                       c->d_scal, thermo_args(c, thermo));
}

void sx_qq_initial_impl(sx_ctx *c, double dt)
{
    hipLaunchKernelGGL(k_qq_initial, dim3(sx_blocks(c->n, 256)), dim3(256), 0, c->stream, c->dstate(), dt);
}

void sx_qq_step_begin_impl(sx_ctx *c, double dt)
{
    NList &nl = c->nl;
    int nb = sx_blocks(c->n, 256);
    sx_ensure_part(c, nb, 1);
    double thresh = nl.rule == SX_REBUILD_REFERENCE ? nl.r_skin_sq : (nl.skin / 2.0) * (nl.skin / 2.0);
    hipLaunchKernelGGL(k_qq_step_begin, dim3(nb), dim3(256), 0, c->stream, c->dstate(), dt, nl.x_last, nl.y_last,
                       nl.z_last, c->d_part, c->d_ticket, thresh, nl.rule, c->d_flag, c->d_scal, c->d_pe_acc);
}

void sx_qq_final_impl(sx_ctx *c, double dt, bool thermo)
{
    int nb = sx_blocks(c->n, 256);
    sx_ensure_part(c, nb, 3);
    hipLaunchKernelGGL(k_qq_final, dim3(nb), dim3(256), 0, c->stream, c->dstate(), dt, c->d_part, c->d_ticket,
                       c->d_scal, thermo_args(c, thermo));
}

static LgvConst lgv_const(sx_ctx *c, double dt, double damp, bool save)
{
    LgvConst k;
    k.dt = dt;
    k.damp = damp;
    k.T = c->h_scal[SX_TARGET_TEMP];
    k.e1 = std::exp(-damp * dt);
    k.e2 = 1 - std::exp(-2.0 * damp * dt);
    k.kk1 = 1.0 - (damp * dt) / 2.0;
    k.kk2 = 1.0 / (1.0 + (damp * dt) / 2.0);
    k.temp_int = (int)(k.T * 1000.0);
    k.save = save ? 1 : 0;
    return k;
}

void sx_lgv_initial_impl(sx_ctx *c, double dt, int step, const int *dstep, double damping, bool constrained)
{
    LgvConst k = lgv_const(c, dt, damping, constrained);
    hipLaunchKernelGGL(k_lgv_initial, dim3(sx_blocks(c->n, 256)), dim3(256), 0, c->stream, c->dstate(), k, step, dstep);
    if (constrained) sx_rattle_impl(c, dt, 0);
}

void sx_lgv_step_begin_impl(sx_ctx *c, double dt, double damping, bool ecp_pp)
{
    LgvConst k = lgv_const(c, dt, damping, false);
    NList &nl = c->nl;
    int nb = sx_blocks(c->n, 256);
    sx_ensure_part(c, nb, 1);
    double thresh = nl.rule == SX_REBUILD_REFERENCE ? nl.r_skin_sq : (nl.skin / 2.0) * (nl.skin / 2.0);
    double *pp = nullptr;
    float *scr = nullptr;
    if (ecp_pp) sx_ecp_records(c, &pp, &scr);
    if (ecp_pp)
        hipLaunchKernelGGL((k_lgv_step_begin<true, 256>), dim3(nb), dim3(256), 0, c->stream, c->dstate(), k, c->d_step,
                           nl.x_last, nl.y_last, nl.z_last, c->d_part, c->d_ticket, thresh, nl.rule, c->d_flag,
                           c->d_scal, c->d_pe_acc, pp, scr);
    else
        hipLaunchKernelGGL((k_lgv_step_begin<false, 256>), dim3(nb), dim3(256), 0, c->stream, c->dstate(), k, c->d_step,
                           nl.x_last, nl.y_last, nl.z_last, c->d_part, c->d_ticket, thresh, nl.rule, c->d_flag,
                           c->d_scal, c->d_pe_acc, pp, scr);
}

void sx_lgv_final_impl(sx_ctx *c, double dt, double damping, bool constrained, bool thermo, int *dstep)
{
    (void)damping;
    // 256-thread workgroups (measured faster than 64 at 32k particles)
    int nb = sx_blocks(c->n, 256);
    sx_ensure_part(c, nb, 4);
    // with constraints the thermo must follow RATTLE (virial_bonded)
    if (c->ecp_apply_in_final && !constrained)
        hipLaunchKernelGGL((k_lgv_final<256, true>), dim3(nb), dim3(256), 0, c->stream, c->dstate(), dt, c->d_part,
                           c->d_ticket, c->d_scal, thermo_args(c, thermo), dstep, sx_ecp_apply_args(c),
                           (const unsigned char *)c->d_live, c->d_pe_acc);
    else
        hipLaunchKernelGGL((k_lgv_final<256, false>), dim3(nb), dim3(256), 0, c->stream, c->dstate(), dt, c->d_part,
                           c->d_ticket, c->d_scal, thermo_args(c, thermo && !constrained), dstep, EcpApplyArgs{},
                           (const unsigned char *)nullptr, (double *)nullptr);
    if (constrained) {
        sx_rattle_impl(c, dt, 1);
        if (thermo) {
            if (c->defer_pe) sx_pe_total_impl(c);
            sx_thermo_impl(c);
        }
    }
}

// the chain's device scratch: doubles (2 x workgroup totals, sums of m v^2,
// partial sums, end state), then the flags (2 x) and the epoch (u64), then
// the ticket; the flags, epoch and ticket start at zero
static const size_t NHT_ND = 2 * NHT_WGMAX * 20 + NHT_WGMAX + NHT_WGMAX * 4 + 8;
static const size_t NHT_SYNC_BYTES = sizeof(unsigned long long) * (2 * NHT_WGMAX + 1) + sizeof(unsigned) * 4;

static NhtBuf nht_buffers(sx_ctx *c)
{
    if (!c->d_nht) {
        HIPCHK(hipMalloc(&c->d_nht, sizeof(double) * NHT_ND + NHT_SYNC_BYTES));
        HIPCHK(hipMemsetAsync(c->d_nht + NHT_ND, 0, NHT_SYNC_BYTES, c->stream));
    }
    NhtBuf B;
    B.T[0] = c->d_nht;
    B.T[1] = B.T[0] + NHT_WGMAX * 20;
    B.msum = B.T[1] + NHT_WGMAX * 20;
    B.part = B.msum + NHT_WGMAX;
    B.s0 = B.part + NHT_WGMAX * 4;
    B.flag = (unsigned long long *)(c->d_nht + NHT_ND);
    B.epoch = B.flag + 2 * NHT_WGMAX;
    B.ticket = (unsigned *)(B.epoch + 1);
    B.chunk = B.nb = 0;
    return B;
}

// after DERR_NHT_SYNC: a call whose workgroups gave up waiting left flags,
// epoch and ticket mid-call
void sx_nht_reset_sync(sx_ctx *c)
{
    if (c->d_nht) HIPCHK(hipMemsetAsync(c->d_nht + NHT_ND, 0, NHT_SYNC_BYTES, c->stream));
}

void sx_nht_impl(sx_ctx *c, double dt, const double *coupling, int phase, bool save)
{
    double ndof = (double)(3 * (c->n / c->nrigid) - 3);
    // the serial walk for small systems, the parallel (DEER) evaluation above
    // a few thousand particles (whose first pass also sets the chain masses)
    if (c->n < 2048) {
        hipLaunchKernelGGL(k_nht_hb, dim3(1), dim3(64), 0, c->stream, c->d_scal, ndof, coupling ? *coupling : 0.0,
                           coupling ? 1 : 0);
        sx_ensure_part(c, 1, 1);
        hipLaunchKernelGGL(k_nht, dim3(1), dim3(NHT_CHUNK), 0, c->stream, c->dstate(), dt, phase, save ? 1 : 0,
                           c->d_scal, ndof, c->d_part);
    } else {
        NhtBuf B = nht_buffers(c);
        B.chunk = (c->n + NHT_CMAX - 1) / NHT_CMAX;
        const int nchunks = (c->n + B.chunk - 1) / B.chunk;
        B.nb = (nchunks + NHT_BLK - 1) / NHT_BLK;
        unsigned long long *st = (unsigned long long *)(c->d_counts + 4);
        // SUCROSE_NHT_PASSES (1..NHT_MAXIT) caps the passes, so that tests can
        // drive the windows from the host
        int passes = NHT_MAXIT;
        if (const char *e = getenv("SUCROSE_NHT_PASSES")) passes = std::max(1, std::min(NHT_MAXIT, atoi(e)));
        hipLaunchKernelGGL(k_nhtp_solve, dim3(B.nb), dim3(NHT_BLK * NHT_DIRS), 0, c->stream, c->dstate(), dt, phase,
                           save ? 1 : 0, c->d_scal, ndof, B, passes, st, coupling ? *coupling : 0.0, coupling ? 1 : 0,
                           c->d_err);
    }
}

extern "C" int sx_nht_stats(sx_ctx *c, long out[3])
{
    return guard(c, [&] {
        long dev[8];
        HIPCHK(hipMemcpyAsync(dev, c->d_counts, sizeof(long) * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        out[0] = dev[4];
        out[1] = dev[5];
        out[2] = dev[6];
    });
}

void sx_rattle_impl(sx_ctx *c, double dt, int phase)
{
    Molecules &m = c->mol;
    if (m.nmol == 0) return;
    sx_ensure_part(c, m.nmol, 1);
    if (m.max_bonds <= SX_WAVE) {
        if (phase == 0)
            hipLaunchKernelGGL(k_rattle_pipe<0>, dim3(m.nmol), dim3(64), 0, c->stream, c->dstate(), dt, m.atom_start,
                               m.atoms, m.bond_start, m.bond_a, m.bond_b, m.bond_len, m.bond_lvl, m.sched, c->d_part,
                               c->d_err);
        else
            hipLaunchKernelGGL(k_rattle_pipe<1>, dim3(m.nmol), dim3(64), 0, c->stream, c->dstate(), dt, m.atom_start,
                               m.atoms, m.bond_start, m.bond_a, m.bond_b, m.bond_len, m.bond_lvl, m.sched, c->d_part,
                               c->d_err);
    } else if (phase == 0)
        hipLaunchKernelGGL(k_rattle<0>, dim3(m.nmol), dim3(64), 0, c->stream, c->dstate(), dt, m.nmol, m.atom_start,
                           m.atoms, m.bond_start, m.bond_a, m.bond_b, m.bond_len, c->d_part, c->d_err);
    else
        hipLaunchKernelGGL(k_rattle<1>, dim3(m.nmol), dim3(64), 0, c->stream, c->dstate(), dt, m.nmol, m.atom_start,
                           m.atoms, m.bond_start, m.bond_a, m.bond_b, m.bond_len, c->d_part, c->d_err);
    reduce_into(c, m.nmol, 1, {SX_VIRIAL_BONDED}, 1);
}


void sx_bonded_impl(sx_ctx *c)
{
    if (c->bonded.empty()) return;
    const double PI_REF = 3.1415926535897932384626433832795;
    for (auto &b : c->bonded) {
        double th0 = PI_REF - (PI_REF / 180.0) * b.angle;
        hipLaunchKernelGGL(k_bonded_terms, dim3(sx_blocks(b.count, 256)), dim3(256), 0, c->stream, c->dstate(), b.kind,
                           b.count, b.idx, b.coeff, th0, std::cos(th0), std::sin(th0), b.out);
    }
    int nb = sx_blocks(c->n, 256);
    sx_ensure_part(c, nb, 1);
    hipLaunchKernelGGL(k_bonded_gather, dim3(nb), dim3(256), 0, c->stream, c->dstate(), c->d_atom_term_start,
                       c->d_atom_term, (const double *const *)c->d_bbufs, c->d_boff, (int)c->bonded.size(),
                       c->d_pe_acc);
    if (!c->defer_pe) sx_pe_total_impl(c);
}

static void rebuild_bonded_index(sx_ctx *c)
{
    int n = c->n;
    std::vector<std::vector<int>> per(n);
    std::vector<const double *> bufs;
    std::vector<long> off;
    long slot0 = 0;
    for (auto &b : c->bonded) {
        std::vector<int> h(4 * (size_t)b.count);
        HIPCHK(hipMemcpy(h.data(), b.idx, sizeof(int) * 4 * b.count, hipMemcpyDeviceToHost));
        int npos = (b.kind <= 2 || b.kind == 5) ? 3 : 4;
        for (int t = 0; t < b.count; ++t)
            for (int p = 0; p < npos; ++p) per[h[4 * t + p]].push_back((int)(((slot0 + t) << 2) | p));
        bufs.push_back(b.out);
        off.push_back(slot0);
        slot0 += b.count;
    }
    std::vector<int> start(n + 1, 0), items;
    for (int i = 0; i < n; ++i) {
        start[i + 1] = start[i] + (int)per[i].size();
        for (int v : per[i]) items.push_back(v);
    }
    if (c->d_atom_term_start) hipFree(c->d_atom_term_start);
    if (c->d_atom_term) hipFree(c->d_atom_term);
    HIPCHK(hipMalloc(&c->d_atom_term_start, sizeof(int) * (n + 1)));
    HIPCHK(hipMalloc(&c->d_atom_term, sizeof(int) * std::max<size_t>(1, items.size())));
    HIPCHK(hipMemcpy(c->d_atom_term_start, start.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    if (!items.empty()) HIPCHK(hipMemcpy(c->d_atom_term, items.data(), sizeof(int) * items.size(), hipMemcpyHostToDevice));
    if (c->d_bbufs) hipFree(c->d_bbufs);
    if (c->d_boff) hipFree(c->d_boff);
    HIPCHK(hipMalloc(&c->d_bbufs, sizeof(double *) * bufs.size()));
    HIPCHK(hipMalloc(&c->d_boff, sizeof(long) * off.size()));
    HIPCHK(hipMemcpy(c->d_bbufs, bufs.data(), sizeof(double *) * bufs.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_boff, off.data(), sizeof(long) * off.size(), hipMemcpyHostToDevice));
}

static void need_temp(sx_ctx *c)
{
    if (!c->have_target_temp) throw sx_error{SX_ERR_NOTEMP, "Error: Target temperature not set."};
}

extern "C" {

int sx_vv_initial_step(sx_ctx *c, double dt) { return guard(c, [&] { need_system(c); sx_vv_initial_impl(c, dt, false); }); }
int sx_vv_final_step(sx_ctx *c, double dt) { return guard(c, [&] { need_system(c); sx_vv_final_impl(c, dt); }); }
int sx_vvc_initial_step(sx_ctx *c, double dt)
{
    return guard(c, [&] { need_system(c); sx_vv_initial_impl(c, dt, true); sx_rattle_impl(c, dt, 0); sx_check_device_error(c); });
}
int sx_vvc_final_step(sx_ctx *c, double dt)
{
    return guard(c, [&] { need_system(c); sx_vv_final_impl(c, dt); sx_rattle_impl(c, dt, 1); sx_check_device_error(c); });
}
int sx_qq_initial_step(sx_ctx *c, double dt) { return guard(c, [&] { need_system(c); sx_qq_initial_impl(c, dt); }); }
int sx_qq_final_step(sx_ctx *c, double dt) { return guard(c, [&] { need_system(c); sx_qq_final_impl(c, dt); }); }

int sx_nht_initial_step(sx_ctx *c, double dt, const double *coupling)
{
    return guard(c, [&] { need_system(c); need_temp(c); sx_nht_impl(c, dt, coupling, 0, false); });
}
int sx_nht_final_step(sx_ctx *c, double dt, const double *coupling)
{
    return guard(c, [&] { need_system(c); sx_nht_impl(c, dt, coupling, 1, false); });
}
int sx_nhtc_initial_step(sx_ctx *c, double dt, const double *coupling)
{
    return guard(c, [&] {
        need_system(c); need_temp(c);
        sx_nht_impl(c, dt, coupling, 0, true);
        sx_rattle_impl(c, dt, 0);
        sx_check_device_error(c);
    });
}
int sx_nhtc_final_step(sx_ctx *c, double dt, const double *coupling)
{
    return guard(c, [&] { need_system(c); sx_nht_impl(c, dt, coupling, 1, false); sx_rattle_impl(c, dt, 1); sx_check_device_error(c); });
}

int sx_lgv_initial_step(sx_ctx *c, double dt, int step, double damping)
{
    return guard(c, [&] { need_system(c); need_temp(c); sx_lgv_initial_impl(c, dt, step, nullptr, damping, false); });
}
int sx_lgv_final_step(sx_ctx *c, double dt, double damping)
{
    return guard(c, [&] { need_system(c); sx_lgv_final_impl(c, dt, damping, false); });
}
int sx_lgvc_initial_step(sx_ctx *c, double dt, int step, double damping)
{
    return guard(c, [&] {
        need_system(c);
        if (c->h_scal[SX_TARGET_TEMP] <= 1e-3) throw sx_error{SX_ERR_NOTEMP, "Error: Target temperature too small."};
        sx_lgv_initial_impl(c, dt, step, nullptr, damping, true);
        sx_check_device_error(c);
    });
}
int sx_lgvc_final_step(sx_ctx *c, double dt, double damping)
{
    return guard(c, [&] { need_system(c); sx_lgv_final_impl(c, dt, damping, true); sx_check_device_error(c); });
}

int sx_apply_constraints_a(sx_ctx *c, double dt)
{
    return guard(c, [&] { need_system(c); sx_rattle_impl(c, dt, 0); sx_check_device_error(c); });
}
int sx_apply_constraints_b(sx_ctx *c, double dt)
{
    return guard(c, [&] { need_system(c); sx_rattle_impl(c, dt, 1); sx_check_device_error(c); });
}

// set_num_constraints + add_constraint; molecules = connected components of the
// bond graph, each keeping its bonds in the global order
int sx_set_constraints(sx_ctx *c, int nc, const int *bbi1, const int *bbj1, const double *len)
{
    return guard(c, [&] {
        need_system(c);
        if (nc < 0) throw sx_error{SX_ERR_ARG, "nc < 0"};
        int n = c->n;
        std::vector<int> parent(n);
        std::iota(parent.begin(), parent.end(), 0);
        auto find = [&](int x) {
            while (parent[x] != x) { parent[x] = parent[parent[x]]; x = parent[x]; }
            return x;
        };
        for (int b = 0; b < nc; ++b) {
            int i = bbi1[b] - 1, j = bbj1[b] - 1;
            if (i < 0 || i >= n || j < 0 || j >= n || i == j)
                throw sx_error{SX_ERR_ARG, "Error: Constraint index does not exist in system"};
            int a = find(i), d = find(j);
            if (a != d) parent[a] = d;
        }
        // molecules in order of their first bond
        std::vector<int> mol_of_root(n, -1), mol_first;
        std::vector<std::vector<int>> mb;
        for (int b = 0; b < nc; ++b) {
            int r = find(bbi1[b] - 1);
            if (mol_of_root[r] < 0) { mol_of_root[r] = (int)mb.size(); mb.emplace_back(); }
            mb[mol_of_root[r]].push_back(b);
        }
        std::vector<int> atom_start{0}, atoms, bond_start{0}, ba, bb, gidx, lvl, sched;
        std::vector<double> bl;
        int max_atoms = 0, max_bonds = 0;
        for (auto &bonds : mb) {
            std::vector<int> loc;  // atoms of the molecule in first-appearance order
            auto local = [&](int g) {
                for (size_t k = 0; k < loc.size(); ++k)
                    if (loc[k] == g) return (int)k;
                loc.push_back(g);
                return (int)loc.size() - 1;
            };
            for (int b : bonds) {
                int la = local(bbi1[b] - 1), lb = local(bbj1[b] - 1);
                ba.push_back(la); bb.push_back(lb); bl.push_back(len[b]); gidx.push_back(b);
            }
            if ((int)loc.size() > SX_WAVE)
                throw sx_error{SX_ERR_ARG, "a constrained molecule has more than 64 atoms"};
            // Pipelined sweeps.  Every bond update reads and writes both of
            // its atoms, so the serial order fixes, for each atom, the order
            // of the updates that touch it.  Bond q of sweep s is given the
            // step lvl[q] + P*s: lvl[q] is one past the step of the previous
            // bond (in the sweep order) that shares an atom with q, and the
            // period P is one past the widest spread of steps over the bonds
            // of one atom, so that an atom's first bond of sweep s+1 comes
            // after its last bond of sweep s.  Each atom then sees its updates
            // in the serial order at strictly increasing steps, and the bonds
            // of one step touch disjoint atoms.
            {
                std::vector<int> first(loc.size(), -1), last(loc.size(), -1);
                int span = 0, period = 1;
                const size_t q0 = ba.size() - bonds.size();
                for (size_t q = q0; q < ba.size(); ++q) {
                    const int a = ba[q], b = bb[q];
                    const int l = std::max(last[a], last[b]) + 1;
                    lvl.push_back(l);
                    span = std::max(span, l);
                    if (first[a] < 0) first[a] = l;
                    if (first[b] < 0) first[b] = l;
                    last[a] = last[b] = l;
                }
                for (size_t k = 0; k < loc.size(); ++k) period = std::max(period, last[k] - first[k] + 1);
                sched.push_back(period);
                sched.push_back(span);
            }
            max_atoms = std::max(max_atoms, (int)loc.size());
            max_bonds = std::max(max_bonds, (int)bonds.size());
            for (int g : loc) atoms.push_back(g);
            atom_start.push_back((int)atoms.size());
            bond_start.push_back((int)ba.size());
        }
        Molecules &m = c->mol;
        void *old[] = {m.atom_start, m.atoms, m.bond_start, m.bond_a, m.bond_b, m.bond_len, m.bond_gidx,
                       m.bond_lvl, m.sched};
        for (void *p : old)
            if (p) hipFree(p);
        m = Molecules();
        m.nmol = (int)mb.size();
        m.max_atoms = max_atoms;
        m.max_bonds = max_bonds;
        auto up = [&](auto &vec, auto **dst) {
            using T = typename std::decay<decltype(vec[0])>::type;
            HIPCHK(hipMalloc(dst, sizeof(T) * std::max<size_t>(1, vec.size())));
            if (!vec.empty()) HIPCHK(hipMemcpy(*dst, vec.data(), sizeof(T) * vec.size(), hipMemcpyHostToDevice));
        };
        up(atom_start, &m.atom_start);
        up(atoms, &m.atoms);
        up(bond_start, &m.bond_start);
        up(ba, &m.bond_a);
        up(bb, &m.bond_b);
        up(bl, &m.bond_len);
        up(gidx, &m.bond_gidx);
        up(lvl, &m.bond_lvl);
        up(sched, &m.sched);
        c->nc = nc;
        c->h_bbi.assign(bbi1, bbi1 + nc);
        c->h_bbj.assign(bbj1, bbj1 + nc);
        c->h_bblen.assign(len, len + nc);
        c->have_graph = false;
    });
}

int sx_set_bonded(sx_ctx *c, int kind, int count, const int *idx1, double coeff, double angle)
{
    return guard(c, [&] {
        need_system(c);
        if (kind < 1 || kind > 6 || count < 0) throw sx_error{SX_ERR_ARG, "bad bonded kind"};
        int npos = (kind <= 2 || kind == 5) ? 3 : 4;
        std::vector<int> h(4 * (size_t)count, 0);
        for (int t = 0; t < count; ++t)
            for (int p = 0; p < npos; ++p) {
                int a = idx1[npos * t + p] - 1;
                if (a < 0 || a >= c->n) throw sx_error{SX_ERR_ARG, "Error: Atom index does not exist in system"};
                h[4 * t + p] = a;
            }
        Bonded b;
        b.kind = kind; b.count = count; b.coeff = coeff; b.angle = angle;
        HIPCHK(hipMalloc(&b.idx, sizeof(int) * std::max(1, 4 * count)));
        HIPCHK(hipMalloc(&b.out, sizeof(double) * std::max(1, 13 * count)));
        if (count) HIPCHK(hipMemcpy(b.idx, h.data(), sizeof(int) * 4 * count, hipMemcpyHostToDevice));
        c->bonded.push_back(b);
        rebuild_bonded_index(c);
        c->have_graph = false;
    });
}

int sx_bonded_forces(sx_ctx *c)
{
    return guard(c, [&] { need_system(c); sx_bonded_impl(c); });
}

}  // extern "C"
