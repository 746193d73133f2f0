// Internal declarations of the gfx950 implementation of sucrose.h.
//
// Numerics: every translation unit is compiled with -ffp-contract=off.  The
// state is FP64 structure-of-arrays in HBM; routines that reproduce a
// discrete decision of the reference (neighbour-list inclusion, cutoff masks,
// the ECP Brent branches, RATTLE tolerances) evaluate the same FP64
// expressions in the same order as the reference, so those decisions agree
// exactly.  FP64 add/mul/div/sqrt are IEEE correctly rounded on gfx950;
// exp/log/sin/cos come from the device math library (within 1 ulp of the
// host libm the reference links).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <exception>
#include <string>
#include <vector>

#include "../../include/sucrose.h"

#define SX_WAVE 64

// ---------------------------------------------------------------------------
// error plumbing
// ---------------------------------------------------------------------------
struct sx_error {
    int code;
    std::string msg;
};

#define HIPCHK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw sx_error{SX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
    } while (0)

// device error word codes (first error wins)
enum { DERR_NONE = 0, DERR_NLIST_CAP = 1, DERR_ROW_CAP = 2, DERR_CONSTRAINT = 3, DERR_ECP_CAP = 4, DERR_NHT_SYNC = 5, DERR_LIST_SYNC = 6 };

// ---------------------------------------------------------------------------
// state passed to kernels by value
// ---------------------------------------------------------------------------
struct DState {
    double *f[SX_NFIELDS];
    int n;
    int periodic;
    double box[3];
};

struct NList {
    bool init = false;
    double cutoff = 0, skin = 0, r_skin_sq = 0;
    int rule = SX_REBUILD_REFERENCE;
    // cell grid
    bool use_cells = false;
    int ncell[3] = {0, 0, 0};
    int ncells = 0;
    // device arrays
    int *cell_of = nullptr, *cell_count = nullptr, *cell_start = nullptr, *cell_items = nullptr;
    int *row_len = nullptr, *row_start = nullptr, *lower = nullptr, *half_start = nullptr, *half_len = nullptr;
    int *nbr = nullptr;
    int *half_i = nullptr, *half_j = nullptr;
    double *x_last = nullptr, *y_last = nullptr, *z_last = nullptr;
    double *cell_pos = nullptr;  // positions in cell order (x y z pad)
    long cap_full = 0, cap_half = 0;
    int max_row = 0;
    int lpr_rows = 16, lpr_half = 16;   // lanes per row of the row and half-index kernels
    int lb_grid = 0;                     // workgroups of the one-launch conditional rebuild (0: stage kernels)
    unsigned *lb_sync = nullptr;         // its barrier words and build epoch
    int *lb_rows = nullptr;              // its scratch rows (n x lb_rc)
    int lb_rc = 0;
    int *lb_bsum = nullptr;              // its per-workgroup row sums
    int *lb_bin_items = nullptr;         // its cell bins (ncells x lb_cap)
    double *lb_bin_pos = nullptr;        // (x y z pad per slot)
    int lb_cap = 0;
    int *lb_meta = nullptr;              // its per-particle row records (int4)
    bool need_half = false;             // half list and pair index kept (ECP paths, export)
    long rebuilds = 0;
    // host copies of the sizes of the last build
    long nnz_full = 0, nnz_half = 0;
};

struct Bonded {
    int kind;        // 1 bend, 2 cosine_bend, 3 dih, 4 cosine_dih, 5 angle_bend, 6 angle_dih
    int count;
    double coeff, angle;
    int *idx = nullptr;   // count * 4 (0-based atoms; 4th unused for bends)
    double *out = nullptr; // count * 13: pe, 4x3 forces
};

struct Molecules {
    int nmol = 0;
    int *atom_start = nullptr, *atoms = nullptr;  // CSR of atoms per molecule
    int *bond_start = nullptr, *bond_a = nullptr, *bond_b = nullptr;  // local atom ids
    double *bond_len = nullptr;
    int *bond_gidx = nullptr;
    // pipelined sweeps: bond q of sweep s runs at step bond_lvl[q] + P*s,
    // sched[2m] = P (the period), sched[2m+1] = the largest bond_lvl
    int *bond_lvl = nullptr, *sched = nullptr;
    int max_atoms = 0, max_bonds = 0;
};

// composite bodies registered for the step driver: placements (centre ->
// satellite, offsets) and the transfer tables (each centre's satellites in
// the order given)
struct Composites {
    int count = 0, ncentres = 0;
    int *centre = nullptr, *sat = nullptr;
    double *off = nullptr;
    int *t_centres = nullptr, *t_start = nullptr, *t_sats = nullptr;
};

struct sx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // system
    int n = 0, nrigid = 1, nc = 0;
    double box[3] = {0, 0, 0};
    int periodic = 0;
    bool have_target_temp = false;
    double *d_fields = nullptr;  // SX_NFIELDS * ld doubles
    long ld = 0;
    double *d_pe_acc = nullptr;  // per-particle potential energy since initialize_forces
    bool defer_pe = false;       // step driver: pot_eng is reduced in the final-step kernel
    bool ecp_pp_by_step = false; // step driver: the initial-step kernel wrote the ECP per-particle records
    bool ecp_apply_in_final = false; // step driver: the final-step kernel applies the ECP pair results
    int *d_segcount = nullptr;   // ECP work-list segment counts
    double *d_scal = nullptr;    // SX_NSCALARS doubles (device copy of the scalars)
    double h_scal[SX_NSCALARS] = {0};
    double target_pres = 0.0;    // set_target_pressure (system.f90:224)
    Composites comp;
    double *d_dip = nullptr;     // dipole table DIP (3 * n) of module dipole
    double *d_rfl = nullptr;     // its reaction field (one vector, shared by all particles)
    int dip_cap = 0;
    bool scal_dirty_host = false; // host copy newer than device
    double *d_part = nullptr;    // reduction partials
    int part_cap = 0;
    int *d_err = nullptr;
    unsigned *d_ticket = nullptr; // last-block reduction counters (self re-arming)
    // host-side metadata the kernels never read: atom types (rtype of
    // module system, -1 until set) and the constraint table as given
    // (1-based), for the reporters
    std::vector<int> h_rtype;
    std::vector<int> h_bbi, h_bbj;
    std::vector<double> h_bblen;
    // constraints
    int *d_bbi = nullptr, *d_bbj = nullptr;
    double *d_bblen = nullptr;
    Molecules mol;
    // neighbour list
    NList nl;
    int *d_flag = nullptr;       // [0] rebuild flag, [1] full nnz, [2] half nnz, [4] ECP work count
    int *d_step = nullptr;       // Langevin step counter for graph replays
    long *d_counts = nullptr;    // device counters (ECP stats etc.)
    // bonded
    std::vector<Bonded> bonded;
    int *d_atom_term_start = nullptr, *d_atom_term = nullptr;  // atom -> (term slot) CSR over all bonded
    void *d_bbufs = nullptr;     // device table of the term buffers
    long *d_boff = nullptr;      // first global slot of each buffer
    bool bonded_index_dirty = true;
    // pair buffers (anisotropic kernels)
    double *d_pairbuf = nullptr;
    unsigned char *d_live = nullptr;
    int *d_work = nullptr;        // compacted pair list
    long pair_cap = 0;
    // per particle: a count and the solved pairs it belongs to, as (partner,
    // half pair), in a line of slots taken by the solve stage; the count is
    // read and cleared again by the apply stage
    int2 *d_ecp_slot = nullptr;
    long ecp_slot_n = 0;
    // parallel Nose-Hoover scratch (maps, guesses, chunk sums, control)
    double *d_nht = nullptr;
    // per-particle ECP precompute
    double *d_ecp_pp = nullptr;   // FP64 records, 24 doubles per particle
    float *d_ecp_scr = nullptr;   // FP32 screen records, ECP_SCR_STRIDE floats per particle
    int ecp_pp_cap = 0;
    // per-particle Gay-Berne factors (GB_TAB doubles per particle)
    double *d_gb_tab = nullptr;
    int gb_tab_cap = 0;
    // FP32 pair terms for LJ, WCA, Debye-Hueckel and gb_ (sx_set_pair_precision)
    bool pair_f32 = false;
    int ecp_solver = SX_ECP_BRENT_REFERENCE;
    long ecp_stats[4] = {0, 0, 0, 0};
    // MLP (learned warm start)
    void *mlp = nullptr;
    // batched replicas (sx_batch_set)
    void *batch = nullptr;
    // step graphs of SX_GRAPH_CHUNK[k] whole steps each (captured on demand;
    // a replay launch has a fixed cost of ~13 us, so long runs replay the
    // longest chunks); all belong to graph_spec while have_graph holds
    hipGraphExec_t graph_exec[4] = {nullptr, nullptr, nullptr, nullptr};
    hipGraph_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
    sx_step_spec graph_spec;
    bool have_graph = false;
    DState dstate() const;
};

struct SxRedIds {
    int id[8];
    int n;
    int accumulate;
};

// Runs fn, converting thrown sx_error into the status code and message.
template <typename F>
static inline int guard(sx_ctx *c, F &&fn)
{
    if (!c) return SX_ERR_ARG;
    try {
        HIPCHK(hipSetDevice(c->device));
        fn();
        return SX_OK;
    } catch (const sx_error &e) {
        c->err = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        c->err = e.what();
        return SX_ERR_HIP;
    }
}

static inline void need_system(sx_ctx *c)
{
    if (c->n <= 0) throw sx_error{SX_ERR_STATE, "Error: System has no atoms. Set number of atoms first."};
}

// ---------------------------------------------------------------------------
// helpers shared across translation units
// ---------------------------------------------------------------------------
void sx_check_device_error(sx_ctx *c);                // syncs, throws on device error
void sx_ensure_part(sx_ctx *c, int nblocks, int width);
void sx_reduce_to_scalar(sx_ctx *c, int nblocks, int width, const int *ids, int nids, int accumulate);
void sx_sync_scalars_to_device(sx_ctx *c);
void sx_sync_scalars_to_host(sx_ctx *c);
void sx_nlist_build_impl(sx_ctx *c, bool conditional);
void sx_nlist_flag_impl(sx_ctx *c);
void sx_nlist_rebox_impl(sx_ctx *c);
// grows the list (contents kept) so that it can take half as many entries
// again as the last build before graph replays
void sx_nlist_headroom(sx_ctx *c);
// device error check after replays; a list that overflowed inside them is
// grown and rebuilt before SX_ERR_CAPACITY is thrown; headroom: then make
// room for the next replays (sx_nlist_headroom)
void sx_nlist_check_after_replay(sx_ctx *c, bool headroom);
void sx_eval_reaction_field_impl(sx_ctx *c);
void sx_composite_place_impl(sx_ctx *c);
void sx_composite_transfer_impl(sx_ctx *c);
void sx_composite_free(sx_ctx *c);
void sx_nlist_want_half(sx_ctx *c);
void sx_forces_impl(sx_ctx *c, int kind, const double *p, int np);
bool sx_forces_two_impl(sx_ctx *c, int k1, const double *p1, int np1, int k2, const double *p2, int np2);
void sx_bonded_impl(sx_ctx *c);
void sx_lgv_initial_impl(sx_ctx *c, double dt, int step, const int *dstep, double damping, bool constrained);
// lgv_initial_step, the rebuild decision and initialize_forces in one launch
// (and the ECP per-particle records of the new orientations when ecp_pp)
void sx_lgv_step_begin_impl(sx_ctx *c, double dt, double damping, bool ecp_pp);
// the ECP per-particle record buffers (allocated on first use)
void sx_ecp_records(sx_ctx *c, double **pp, float **scr);
// thermo = also run calculate_thermo in the same kernel; dstep = device step
// counter to advance (graph replays)
void sx_lgv_final_impl(sx_ctx *c, double dt, double damping, bool constrained, bool thermo = false,
                       int *dstep = nullptr);
void sx_vv_initial_impl(sx_ctx *c, double dt, bool save);
void sx_vv_final_impl(sx_ctx *c, double dt, bool thermo = false);
void sx_qq_initial_impl(sx_ctx *c, double dt);
void sx_qq_step_begin_impl(sx_ctx *c, double dt);   // qq_initial + rebuild flag + initialize_forces
void sx_nht_reset_sync(sx_ctx *c);   // clears the chain's call flags after DERR_NHT_SYNC
void sx_nlist_reset_sync(sx_ctx *c);   // clears the one-launch rebuild's flags after DERR_LIST_SYNC
void sx_qq_final_impl(sx_ctx *c, double dt, bool thermo = false);
void sx_nht_impl(sx_ctx *c, double dt, const double *coupling, int phase, bool save);
void sx_rattle_impl(sx_ctx *c, double dt, int phase);
void sx_thermo_impl(sx_ctx *c);
void sx_init_forces_impl(sx_ctx *c);
// pot_eng = sum of the per-particle energies accumulated since initialize_forces
void sx_pe_total_impl(sx_ctx *c);
void sx_ecp_mlp_free(sx_ctx *c);
void sx_batch_free(sx_ctx *c);

static inline int sx_blocks(long n, int bs) { return (int)((n + bs - 1) / bs); }

// ---------------------------------------------------------------------------
// device helpers (FP64, reference evaluation order)
// ---------------------------------------------------------------------------
// minimum image, d - anint(d/L)*L.  For |d| < L/2 the correctly rounded
// quotient is at most the double just below 1/2, so anint gives a signed zero
// and the result is d + 0.0 (which also turns -0.0 into +0.0 as the full
// expression does): the same bits without the division.
__device__ __forceinline__ double sx_mimg(double d, double L)
{
    if (fabs(d) < 0.5 * L) return d + 0.0;
    return d - round(d / L) * L;
}

// the same value always through the division, for callers (all pairs of a
// small box) where half the components lie beyond L/2 and a wavefront would
// take both paths
__device__ __forceinline__ double sx_mimg_div(double d, double L) { return d - round(d / L) * L; }

// Same value as sx_mimg whenever |d| is not within a few ulp of an odd
// multiple of L/2: round(d*invL) == round(d/L) there, and the subtraction
// d - k*L is identical.  Callers use it only for pairs whose distance is
// compared against a radius below L/2, where a half-box component excludes
// the pair under either image.
__device__ __forceinline__ double sx_mimg_fast(double d, double L, double invL) { return d - round(d * invL) * L; }

// Deterministic one-pass block reduction of W doubles per thread.  Every
// block writes its tree-reduced partial; the last block to finish (ticket
// counter) sums the partials in block order and stores (or adds) the totals
// into scal[ids[q]], then re-arms the counter for the next launch.
// blockDim.x must be 256.
// Returns true in the last block, where tot[] (valid in thread 0) holds the
// grid totals.
// Thread 0 of a workgroup publishes its W partial values and takes a ticket;
// true in the last workgroup to do so.  The partials are agent-scope atomic
// stores, completed (s_waitcnt) before the ticket increment, and the last
// workgroup reads them with agent-scope atomic loads issued after its ticket
// returned: coherent across the XCDs' L2 caches without the L2 write-back and
// invalidation of __threadfence().
template <int W>
__device__ __forceinline__ bool sx_publish_take_ticket(double *part, const double (&v)[W], unsigned *ticket)
{
#pragma unroll
    for (int q = 0; q < W; ++q) __hip_atomic_store(&part[blockIdx.x * W + q], v[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    return __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1;
}

// re-arms a ticket for the next launch (in the last workgroup)
__device__ __forceinline__ void sx_ticket_reset(unsigned *ticket)
{
    __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int W, int BS = 256>
__device__ __forceinline__ bool sx_block_reduce_last(double (&v)[W], double *part, unsigned *ticket, double (&tot)[W])
{
    __shared__ double sh[W][BS];
    __shared__ bool last;
    int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < W; ++q) sh[q][t] = v[q];
    __syncthreads();
    for (int s = BS / 2; s > 0; s >>= 1) {
        if (t < s) {
#pragma unroll
            for (int q = 0; q < W; ++q) sh[q][t] = sh[q][t] + sh[q][t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
        double p[W];
#pragma unroll
        for (int q = 0; q < W; ++q) p[q] = sh[q][0];
        last = sx_publish_take_ticket<W>(part, p, ticket);
    }
    __syncthreads();
    if (!last) return false;
    int nb = gridDim.x;
#pragma unroll
    for (int q = 0; q < W; ++q) {
        double a = 0.0;
        for (int b = t; b < nb; b += BS)
            a = a + __hip_atomic_load(&part[b * W + q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        sh[q][t] = a;
    }
    __syncthreads();
    for (int s = BS / 2; s > 0; s >>= 1) {
        if (t < s) {
#pragma unroll
            for (int q = 0; q < W; ++q) sh[q][t] = sh[q][t] + sh[q][t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
#pragma unroll
        for (int q = 0; q < W; ++q) tot[q] = sh[q][0];
        sx_ticket_reset(ticket);
    }
    return true;
}

template <int W>
__device__ __forceinline__ void sx_reduce_last_block(double (&v)[W], double *part, unsigned *ticket, double *scal,
                                                     SxRedIds ids)
{
    double tot[W];
    if (!sx_block_reduce_last<W>(v, part, ticket, tot)) return;
    if (threadIdx.x == 0) {
        for (int q = 0; q < W && q < ids.n; ++q) {
            if (ids.accumulate) scal[ids.id[q]] = scal[ids.id[q]] + tot[q];
            else scal[ids.id[q]] = tot[q];
        }
    }
}

__host__ __device__ static inline SxRedIds sx_ids1(int a, int accumulate)
{
    SxRedIds r;
    r.id[0] = a; r.n = 1; r.accumulate = accumulate;
    return r;
}

__host__ __device__ static inline SxRedIds sx_ids2(int a, int b, int accumulate)
{
    SxRedIds r;
    r.id[0] = a; r.id[1] = b; r.n = 2; r.accumulate = accumulate;
    return r;
}

// Philox4x32-10 with the reference's constants and fixed key (xmath.f90:5-76)
__device__ __forceinline__ void sx_philox(uint32_t c[4])
{
    uint32_t k0 = 0x6D2B79F5u, k1 = 0x1B873593u;
#pragma unroll
    for (int r = 1; r <= 10; ++r) {
        uint32_t lo0 = 0xD2511F53u * c[0];
        uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
        uint32_t lo1 = 0xCD9E8D57u * c[2];
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
        uint32_t n0 = hi1 ^ (c[1] ^ k0);
        uint32_t n2 = hi0 ^ (c[3] ^ k1);
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        if (r < 10) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    }
}

// philox_normal_3seq (xmath.f90:78-113)
__device__ __forceinline__ void sx_normal3(int c1, int c2, int c3, int c4, double &n1, double &n2, double &n3)
{
    uint32_t c[4] = {(uint32_t)c1, (uint32_t)c2, (uint32_t)c3, (uint32_t)c4};
    sx_philox(c);
    const double s = 1.0 / 4294967296.0;
    double u1 = (double)c[0] * s, u2 = (double)c[1] * s, u3 = (double)c[2] * s, u4 = (double)c[3] * s;
    u1 = fmax(u1, 1.0e-35);
    u3 = fmax(u3, 1.0e-35);
    double r1 = sqrt(-2.0 * log(u1));
    double th1 = 6.283185307179586 * u2;
    double r2 = sqrt(-2.0 * log(u3));
    double th2 = 6.283185307179586 * u4;
    n1 = r1 * cos(th1);
    n2 = r1 * sin(th1);
    n3 = r2 * cos(th2);
}

// 3x3 SPD solve in the reference's order (xmath.f90:281-304); a is symmetric,
// given by its lower triangle a11 a21 a31 a22 a32 a33.
__device__ __forceinline__ void sx_chol_solve(double a11, double a21, double a31, double a22, double a32,
                                              double a33, const double b[3], double out[3])
{
    double inv_l11 = 1.0 / sqrt(a11);
    double l21 = a21 * inv_l11;
    double l31 = a31 * inv_l11;
    double inv_l22 = 1.0 / sqrt(a22 - l21 * l21);
    double l32 = (a32 - l31 * l21) * inv_l22;
    double inv_l33 = 1.0 / sqrt(a33 - l31 * l31 - l32 * l32);
    double y1 = b[0] * inv_l11;
    double y2 = (b[1] - l21 * y1) * inv_l22;
    double y3 = (b[2] - l31 * y1 - l32 * y2) * inv_l33;
    out[2] = y3 * inv_l33;
    out[1] = (y2 - l32 * out[2]) * inv_l22;
    out[0] = (y1 - l21 * out[1] - l31 * out[2]) * inv_l11;
}
