// Context, state, scalars, reductions, thermo and set-up routines of sucrose.h.
#include "sx_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

static thread_local std::string g_create_error;

DState sx_ctx::dstate() const
{
    DState s;
    for (int k = 0; k < SX_NFIELDS; ++k) s.f[k] = d_fields + (long)k * ld;
    s.n = n;
    s.periodic = periodic;
    s.box[0] = box[0]; s.box[1] = box[1]; s.box[2] = box[2];
    return s;
}

// ---------------------------------------------------------------------------
// reductions: per-block partials (fixed tree order) then one ordered pass
// ---------------------------------------------------------------------------
// one block of 256 threads: strided partial sums then a fixed tree, per quantity
__global__ void __launch_bounds__(256) k_finish_reduce(const double *part, int nblocks, int width, SxRedIds ids,
                                                       double *scal)
{
    __shared__ double sh[256];
    int t = threadIdx.x;
    for (int q = 0; q < ids.n; ++q) {
        double s = 0.0;
        for (int b = t; b < nblocks; b += 256) s += part[(long)b * width + q];
        sh[t] = s;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (t < o) sh[t] = sh[t] + sh[t + o];
            __syncthreads();
        }
        if (t == 0) {
            if (ids.accumulate) scal[ids.id[q]] = scal[ids.id[q]] + sh[0];
            else scal[ids.id[q]] = sh[0];
        }
        __syncthreads();
    }
}

void sx_ensure_part(sx_ctx *c, int nblocks, int width)
{
    int need = nblocks * width + 64;
    if (need > c->part_cap) {
        if (c->d_part) HIPCHK(hipFree(c->d_part));
        HIPCHK(hipMalloc(&c->d_part, sizeof(double) * need));
        c->part_cap = need;
    }
}

void sx_reduce_to_scalar(sx_ctx *c, int nblocks, int width, const int *ids, int nids, int accumulate)
{
    SxRedIds r;
    r.n = nids;
    r.accumulate = accumulate;
    for (int q = 0; q < nids && q < 8; ++q) r.id[q] = ids[q];
    hipLaunchKernelGGL(k_finish_reduce, dim3(1), dim3(256), 0, c->stream, c->d_part, nblocks, width, r, c->d_scal);
}

void sx_sync_scalars_to_device(sx_ctx *c)
{
    HIPCHK(hipMemcpyAsync(c->d_scal, c->h_scal, sizeof(double) * SX_NSCALARS, hipMemcpyHostToDevice,
                          c->stream));
}

void sx_sync_scalars_to_host(sx_ctx *c)
{
    HIPCHK(hipMemcpyAsync(c->h_scal, c->d_scal, sizeof(double) * SX_NSCALARS, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
}

void sx_check_device_error(sx_ctx *c)
{
    int e = 0;
    HIPCHK(hipMemcpyAsync(&e, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (e != DERR_NONE) {
        HIPCHK(hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream));
        switch (e) {
        case DERR_NLIST_CAP: throw sx_error{SX_ERR_CAPACITY, "neighbour list capacity exceeded"};
        case DERR_ROW_CAP: throw sx_error{SX_ERR_CAPACITY, "neighbour row longer than the sort buffer"};
        case DERR_CONSTRAINT: throw sx_error{SX_ERR_CONSTRAINT, "Constraint failure (RATTLE)"};
        case DERR_ECP_CAP: throw sx_error{SX_ERR_CAPACITY, "ECP work list capacity exceeded"};
        case DERR_LIST_SYNC:
            sx_nlist_reset_sync(c);
            throw sx_error{SX_ERR_HIP, "neighbour list rebuild: the workgroups of one build were not resident "
                                       "together (the list is invalid; rebuild it)"};
        case DERR_NHT_SYNC:
            sx_nht_reset_sync(c);
            throw sx_error{SX_ERR_HIP, "Nose-Hoover chain: the workgroups of one call were not resident together "
                                       "(the call's results are invalid)"};
        default: throw sx_error{SX_ERR_HIP, "device error " + std::to_string(e)};
        }
    }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void k_fill(double *p, long n, double v)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// initialize_forces (system.f90:423-436): F, T and the accumulators
__global__ void k_init_forces(DState s, double *scal, double *pe_acc)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        scal[SX_POT_ENG] = 0.0; scal[SX_KIN_ENG] = 0.0; scal[SX_KIN_ROT] = 0.0; scal[SX_TEMP] = 0.0;
        scal[SX_PRESSURE] = 0.0; scal[SX_VIRIAL] = 0.0; scal[SX_VIRIAL_BONDED] = 0.0;
    }
    if (i >= s.n) return;
    s.f[SX_FX][i] = 0.0; s.f[SX_FY][i] = 0.0; s.f[SX_FZ][i] = 0.0;
    s.f[SX_TX][i] = 0.0; s.f[SX_TY][i] = 0.0; s.f[SX_TZ][i] = 0.0;
    pe_acc[i] = 0.0;
}

__global__ void __launch_bounds__(256) k_pe_total(int n, const double *pe_acc, double *part, unsigned *ticket,
                                                  double *scal)
{
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[0] = v[0] + pe_acc[i];
    sx_reduce_last_block<1>(v, part, ticket, scal, sx_ids1(SX_POT_ENG, 0));
}

// defaults of initialize_system (system.f90:586-682)
__global__ void k_init_defaults(DState s)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    double **f = s.f;
    if (f[SX_CHARGE][i] == -1 || f[SX_CHARGE][i] < 0) f[SX_CHARGE][i] = 0.0;
    if (f[SX_SIG_X][i] == -1 || f[SX_SIG_X][i] < 0) { f[SX_SIG_X][i] = 1.0; f[SX_SIG_Y][i] = 1.0; f[SX_SIG_Z][i] = 1.0; }
    if (f[SX_EPS_X][i] == -1 || f[SX_EPS_X][i] < 0) { f[SX_EPS_X][i] = 1.0; f[SX_EPS_Y][i] = 1.0; f[SX_EPS_Z][i] = 1.0; }
    bool shape_inertia = false;
    if (f[SX_MASS][i] == -1 || f[SX_MASS][i] < 0) {
        f[SX_MASS][i] = 1.0; f[SX_INV_MASS][i] = 1.0; shape_inertia = true;
    } else if (f[SX_MASS][i] == 0) {
        f[SX_INV_MASS][i] = 0.0;
        f[SX_IXX][i] = 0.0; f[SX_IYY][i] = 0.0; f[SX_IZZ][i] = 0.0;
        f[SX_IXX_INV][i] = 0.0; f[SX_IYY_INV][i] = 0.0; f[SX_IZZ_INV][i] = 0.0;
        return;
    } else {
        f[SX_INV_MASS][i] = 1.0 / f[SX_MASS][i]; shape_inertia = true;
    }
    if (shape_inertia) {
        double sx = f[SX_SIG_X][i], sy = f[SX_SIG_Y][i], sz = f[SX_SIG_Z][i];
        if (f[SX_IXX][i] < 0 || f[SX_IYY][i] < 0 || f[SX_IZZ][i] < 0) {
            f[SX_IXX][i] = (sy * sy + sz * sz) / 20.0;
            f[SX_IYY][i] = (sx * sx + sz * sz) / 20.0;
            f[SX_IZZ][i] = (sx * sx + sy * sy) / 20.0;
            f[SX_IXX_INV][i] = 1.0 / f[SX_IXX][i];
            f[SX_IYY_INV][i] = 1.0 / f[SX_IYY][i];
            f[SX_IZZ_INV][i] = 1.0 / f[SX_IZZ][i];
        } else if (f[SX_IXX][i] > 0 && f[SX_IYY][i] > 0 && f[SX_IZZ][i] > 0) {
            f[SX_IXX_INV][i] = 1.0 / f[SX_IXX][i];
            f[SX_IYY_INV][i] = 1.0 / f[SX_IYY][i];
            f[SX_IZZ_INV][i] = 1.0 / f[SX_IZZ][i];
        } else {
            f[SX_IXX_INV][i] = 0.0; f[SX_IYY_INV][i] = 0.0; f[SX_IZZ_INV][i] = 0.0;
        }
    }
}

// block tree-sum of up to W quantities into part[blockIdx.x*W + q]
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

// calculate_thermo (system.f90:725-741): virial = SUM(RX*FX + RY*FY + RZ*FZ)
// (+ virial_bonded), temp = 2 (KE + KErot) / dof, pressure; one pass
__global__ void __launch_bounds__(256) k_thermo(DState s, double *part, unsigned *ticket, double *scal, int nc,
                                                int nrigid, double volume)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (i < s.n)
        v[0] = s.f[SX_RX][i] * s.f[SX_FX][i] + s.f[SX_RY][i] * s.f[SX_FY][i] + s.f[SX_RZ][i] * s.f[SX_FZ][i];
    double tot[1];
    if (!sx_block_reduce_last<1>(v, part, ticket, tot) || threadIdx.x != 0) return;
    double w = tot[0];
    double kin_tot = scal[SX_KIN_ENG] + scal[SX_KIN_ROT];
    scal[SX_TEMP] = 2.0 * kin_tot / scal[SX_DOF];
    if (nc > 0 && scal[SX_VIRIAL_BONDED] >= 0.0) w = w + scal[SX_VIRIAL_BONDED];
    scal[SX_VIRIAL] = w;
    scal[SX_PRESSURE] = (scal[SX_DENSITY] * nrigid * scal[SX_TEMP]) + (scal[SX_VIRIAL] / (3.0 * volume));
}

// set_velocity_to_temperature (system.f90:362-421)
__global__ void k_vel_normals(DState s, int temp_int, int seed)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    double a, b, c;
    sx_normal3(i + 1, temp_int, seed, 42, a, b, c);
    s.f[SX_VX][i] = a; s.f[SX_VY][i] = b; s.f[SX_VZ][i] = c;
}

__global__ void __launch_bounds__(256) k_vsum_part(DState s, double *part)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    if (i < s.n) { v[0] = s.f[SX_VX][i]; v[1] = s.f[SX_VY][i]; v[2] = s.f[SX_VZ][i]; }
    block_sum_store<3>(v, part);
}

__global__ void __launch_bounds__(256) k_vsq_part(DState s, double *part)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (i < s.n) {
        double x = s.f[SX_VX][i], y = s.f[SX_VY][i], z = s.f[SX_VZ][i];
        v[0] = (x * x + y * y) + z * z;
    }
    block_sum_store<1>(v, part);
}

__global__ void k_vel_com_sub(DState s, const double *part, int nblocks)
{
    __shared__ double com[3];
    if (threadIdx.x < 3) {
        double t = 0.0;
        for (int b = 0; b < nblocks; ++b) t += part[b * 3 + threadIdx.x];
        com[threadIdx.x] = t / (double)s.n;
    }
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
        s.f[SX_VX][i] = s.f[SX_VX][i] - com[0];
        s.f[SX_VY][i] = s.f[SX_VY][i] - com[1];
        s.f[SX_VZ][i] = s.f[SX_VZ][i] - com[2];
    }
}

__global__ void k_vel_scale(DState s, const double *part, int nblocks, double numer)
{
    __shared__ double factor;
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int b = 0; b < nblocks; ++b) t += part[b];
        factor = sqrt(numer / t);
    }
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
        s.f[SX_VX][i] = s.f[SX_VX][i] * factor;
        s.f[SX_VY][i] = s.f[SX_VY][i] * factor;
        s.f[SX_VZ][i] = s.f[SX_VZ][i] * factor;
    }
}

__global__ void k_philox_normals(int n, int c2, int c3, int c4, double *out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a, b, c;
    sx_normal3(i + 1, c2, c3, c4, a, b, c);
    out[3 * i] = a; out[3 * i + 1] = b; out[3 * i + 2] = c;
}

// Device scratch of the allocation-free entry point sx_philox_normals_noalloc:
// SX_SCRATCH_DOUBLES doubles (4 MiB) on every GPU the library runs on, held
// until the process ends.  Host code never uses the array's name as an
// address: sx_create (with its device current) asks hipGetSymbolAddress for
// that device's copy once, and every use goes through g_scratch_dev[device].
#define SX_SCRATCH_DOUBLES (1L << 19)
#define SX_SCRATCH_DEVICES 64
static __device__ __align__(256) double scratch[SX_SCRATCH_DOUBLES];
static std::mutex g_scratch_mu;
static double *g_scratch_dev[SX_SCRATCH_DEVICES];
static std::string g_scratch_err[SX_SCRATCH_DEVICES];

// the calling thread's current device must be `device`
static void scratch_setup(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= SX_SCRATCH_DEVICES || g_scratch_dev[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return;
    }
    g_scratch_dev[device] = static_cast<double *>(p);
}

static double *scratch_address(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= SX_SCRATCH_DEVICES)
        throw sx_error{SX_ERR_ARG, "no allocation-free scratch for device indices beyond 63"};
    if (!g_scratch_dev[device])
        throw sx_error{SX_ERR_HIP, "allocation-free scratch unavailable on this device: " + g_scratch_err[device]};
    return g_scratch_dev[device];
}

// ---------------------------------------------------------------------------
// host implementations
// ---------------------------------------------------------------------------
void sx_init_forces_impl(sx_ctx *c)
{
    hipLaunchKernelGGL(k_init_forces, dim3(sx_blocks(std::max(c->n, 1), 256)), dim3(256), 0, c->stream, c->dstate(),
                       c->d_scal, c->d_pe_acc);
}

void sx_pe_total_impl(sx_ctx *c)
{
    int nb = std::min(64, sx_blocks(std::max(c->n, 1), 256));
    sx_ensure_part(c, nb, 1);
    hipLaunchKernelGGL(k_pe_total, dim3(nb), dim3(256), 0, c->stream, c->n, c->d_pe_acc, c->d_part, c->d_ticket,
                       c->d_scal);
}

void sx_thermo_impl(sx_ctx *c)
{
    int nb = sx_blocks(c->n, 256);
    sx_ensure_part(c, nb, 1);
    double volume = c->box[0] * c->box[1] * c->box[2];
    hipLaunchKernelGGL(k_thermo, dim3(nb), dim3(256), 0, c->stream, c->dstate(), c->d_part, c->d_ticket, c->d_scal,
                       c->nc, c->nrigid, volume);
}

static void free_system(sx_ctx *c)
{
    if (c->d_fields) hipFree(c->d_fields);
    if (c->d_pe_acc) hipFree(c->d_pe_acc);
    c->d_fields = nullptr;
    c->d_pe_acc = nullptr;
}

extern "C" {

int sx_version(void) { return 10000; }

int sx_create(int device, sx_ctx **out)
{
    if (!out) return SX_ERR_ARG;
    *out = nullptr;
    sx_ctx *c = new sx_ctx();
    c->device = device;
    try {
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) throw sx_error{SX_ERR_ARG, "no such device"};
        HIPCHK(hipSetDevice(device));
        scratch_setup(device);   // once per device: the address of its scratch array
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIPCHK(hipMalloc(&c->d_scal, sizeof(double) * SX_NSCALARS));
        HIPCHK(hipMalloc(&c->d_err, sizeof(int) * 4));
        HIPCHK(hipMemsetAsync(c->d_err, 0, sizeof(int) * 4, c->stream));
        HIPCHK(hipMalloc(&c->d_ticket, sizeof(unsigned) * 16));
        HIPCHK(hipMemsetAsync(c->d_ticket, 0, sizeof(unsigned) * 16, c->stream));
        HIPCHK(hipMalloc(&c->d_segcount, sizeof(int) * 4096));
        HIPCHK(hipMemsetAsync(c->d_segcount, 0, sizeof(int) * 4096, c->stream));
        HIPCHK(hipMalloc(&c->d_flag, sizeof(int) * 8));
        HIPCHK(hipMemsetAsync(c->d_flag, 0, sizeof(int) * 8, c->stream));
        HIPCHK(hipMalloc(&c->d_counts, sizeof(long) * 8));
        HIPCHK(hipMemsetAsync(c->d_counts, 0, sizeof(long) * 8, c->stream));
        // reference initial values of the nose_hoover state (integrators.f90:230)
        c->h_scal[SX_HB0] = c->h_scal[SX_HB1] = c->h_scal[SX_HB2] = 1.0;
        sx_sync_scalars_to_device(c);
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const sx_error &e) {
        g_create_error = e.msg;
        delete c;
        return e.code;
    }
    *out = c;
    return SX_OK;
}

int sx_destroy(sx_ctx *c)
{
    if (!c) return SX_ERR_ARG;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    free_system(c);
    void *ptrs[] = {c->d_scal, c->d_part, c->d_err, c->d_ticket, c->d_segcount, c->d_flag, c->d_counts, c->d_bbi, c->d_bbj, c->d_bblen,
                    c->nl.cell_of, c->nl.cell_count, c->nl.cell_start, c->nl.cell_items, c->nl.row_len,
                    c->nl.row_start, c->nl.lower, c->nl.half_start, c->nl.nbr, c->nl.half_i,
                    c->nl.half_j, c->nl.x_last, c->nl.y_last, c->nl.z_last, c->d_pairbuf, c->d_live, c->d_work,
                    c->d_ecp_slot, c->d_nht, c->d_dip, c->d_rfl, c->d_ecp_pp, c->d_ecp_scr, c->d_atom_term_start, c->d_atom_term, c->d_bbufs, c->d_boff, c->d_step, c->mol.atom_start, c->mol.atoms,
                    c->mol.bond_start, c->mol.bond_a, c->mol.bond_b, c->mol.bond_len, c->mol.bond_gidx, c->mol.bond_lvl, c->mol.sched};
    for (void *p : ptrs)
        if (p) hipFree(p);
    for (auto &b : c->bonded) {
        if (b.idx) hipFree(b.idx);
        if (b.out) hipFree(b.out);
    }
    if (c->d_gb_tab) hipFree(c->d_gb_tab);
    sx_ecp_mlp_free(c);
    sx_batch_free(c);
    sx_composite_free(c);
    for (int k = 0; k < 4; ++k) {
        if (c->graph_exec[k]) hipGraphExecDestroy(c->graph_exec[k]);
        if (c->graph[k]) hipGraphDestroy(c->graph[k]);
    }
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return SX_OK;
}

const char *sx_last_error(const sx_ctx *c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int sx_synchronize(sx_ctx *c)
{
    return guard(c, [&] { sx_check_device_error(c); });
}

int sx_set_num_atoms(sx_ctx *c, int num_atoms, int rigid_atoms)
{
    return guard(c, [&] {
        if (num_atoms <= 0 || rigid_atoms <= 0) throw sx_error{SX_ERR_ARG, "num_atoms and rigid_atoms must be > 0"};
        free_system(c);
        // registered bonded terms name atoms of the previous system
        for (auto &b : c->bonded) {
            if (b.idx) hipFree(b.idx);
            if (b.out) hipFree(b.out);
        }
        c->bonded.clear();
        c->have_graph = false;
        c->nrigid = rigid_atoms;
        c->n = num_atoms * rigid_atoms;
        c->nc = 0;
        c->h_rtype.assign(c->n, -1);
        c->h_bbi.clear(); c->h_bbj.clear(); c->h_bblen.clear();
        c->ld = ((long)c->n + 63) / 64 * 64;
        HIPCHK(hipMalloc(&c->d_fields, sizeof(double) * SX_NFIELDS * c->ld));
        HIPCHK(hipMalloc(&c->d_pe_acc, sizeof(double) * c->ld));
        hipLaunchKernelGGL(k_fill, dim3(sx_blocks(c->ld, 256)), dim3(256), 0, c->stream, c->d_pe_acc, c->ld, 0.0);
        // sentinels of set_num_atoms (system.f90:74-85)
        const double init[SX_NFIELDS] = {
            0, 0, 0, 0, 0, 0, 0, 0, 0,   // r v f
            1, 0, 0, 0, 0, 0, 0, 0, 0, 0, // q L T
            -1, -1, 0, -1, -1, -1,        // mass inv_mass charge sig
            -1, -1, -1, -1, -1, -1,       // eps I
            -1, -1, -1, 0, 0, 0};         // I_inv r_old
        for (int k = 0; k < SX_NFIELDS; ++k)
            hipLaunchKernelGGL(k_fill, dim3(sx_blocks(c->ld, 256)), dim3(256), 0, c->stream,
                               c->d_fields + k * c->ld, c->ld, init[k]);
        for (int k = 0; k < SX_NSCALARS; ++k)
            if (k < SX_HB0) c->h_scal[k] = 0.0;
        c->h_scal[SX_DOF] = 6.0 * c->n;
        sx_sync_scalars_to_device(c);
        c->nl.init = false;
        c->have_graph = false;
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int sx_set_box(sx_ctx *c, const double box[3], int periodic)
{
    return guard(c, [&] {
        need_system(c);
        if (!box || box[0] == 0.0 || box[1] == 0.0 || box[2] == 0.0)
            throw sx_error{SX_ERR_ARG, "Error: Box dimensions cannot be zero."};
        c->box[0] = box[0]; c->box[1] = box[1]; c->box[2] = box[2];
        c->periodic = periodic ? 1 : 0;
        sx_sync_scalars_to_host(c);
        c->h_scal[SX_DENSITY] = (double)(c->n / c->nrigid) / (box[0] * box[1] * box[2]);
        sx_sync_scalars_to_device(c);
        c->have_graph = false;
    });
}

// set_density (system.f90:167-183): a cubic box of (N / Nrigid) / dens
static void set_density_impl(sx_ctx *c, double dens)
{
    if (dens == 0.0) throw sx_error{SX_ERR_ARG, "Error: Density cannot be zero."};
    double vol = (double)(c->n / c->nrigid) / dens;
    double box_len = std::pow(vol, 1.0 / 3.0);
    c->box[0] = c->box[1] = c->box[2] = box_len;
    sx_sync_scalars_to_host(c);
    c->h_scal[SX_DENSITY] = dens;
    sx_sync_scalars_to_device(c);
    c->have_graph = false;
}

int sx_set_density(sx_ctx *c, double dens)
{
    return guard(c, [&] { need_system(c); set_density_impl(c, dens); });
}

int sx_generate_fcc_lattice(sx_ctx *c, int unit_cells, double dens, int rigid_atoms)
{
    return guard(c, [&] {
        if (unit_cells <= 0) throw sx_error{SX_ERR_ARG, "unit_cells must be > 0"};
        int num_atoms = 4 * unit_cells * unit_cells * unit_cells;
        if (c->n <= 0) {
            int rc = sx_set_num_atoms(c, num_atoms, rigid_atoms > 0 ? rigid_atoms : 1);
            if (rc != SX_OK) throw sx_error{rc, c->err};
        } else if (c->n < num_atoms) {
            throw sx_error{SX_ERR_ARG, "Error: FCC lattice of " + std::to_string(num_atoms) +
                                           " atoms cannot be generated for system of " + std::to_string(c->n) +
                                           " atoms."};
        }
        set_density_impl(c, dens);
        double cell = c->box[0] / (double)unit_cells;
        double half_cell = cell / 2.0;
        double rroot3 = 1.0 / std::sqrt(3.0);
        double den = std::sqrt(1.0 - rroot3 * rroot3);
        double sp = std::sqrt((1.0 + rroot3) / 2.0), sm = std::sqrt((1.0 - rroot3) / 2.0);
        // the four basis sites and their orientations (system.f90:466-492)
        const double bx[4] = {0.0, half_cell, 0.0, half_cell}, by[4] = {0.0, half_cell, half_cell, 0.0},
                     bz[4] = {0.0, 0.0, half_cell, half_cell};
        const double qw[4] = {sp, sm, sm, sp};
        const double qx[4] = {sm * (rroot3 / den), sp * (-rroot3 / den), sp * (rroot3 / den), sm * (-rroot3 / den)};
        const double qy[4] = {sm * (-rroot3 / den), sp * (-rroot3 / den), sp * (rroot3 / den), sm * (rroot3 / den)};
        std::vector<double> f[7];
        for (auto &v : f) v.resize(num_atoms);
        int m = 0;
        for (int I = 1; I <= unit_cells; ++I)
            for (int J = 1; J <= unit_cells; ++J)
                for (int K = 1; K <= unit_cells; ++K) {
                    for (int r = 0; r < 4; ++r) {
                        f[0][m + r] = bx[r] + cell * (double)(I - 1);
                        f[1][m + r] = by[r] + cell * (double)(J - 1);
                        f[2][m + r] = bz[r] + cell * (double)(K - 1);
                        f[3][m + r] = qw[r];
                        f[4][m + r] = qx[r];
                        f[5][m + r] = qy[r];
                        f[6][m + r] = 0.0;
                    }
                    m += 4;
                }
        for (int i = 0; i < num_atoms; ++i) {   // shifted by -box/2 (:518-520)
            f[0][i] = f[0][i] - c->box[0] / 2.0;
            f[1][i] = f[1][i] - c->box[1] / 2.0;
            f[2][i] = f[2][i] - c->box[2] / 2.0;
        }
        const int fl[7] = {SX_RX, SX_RY, SX_RZ, SX_QW, SX_QX, SX_QY, SX_QZ};
        for (int k = 0; k < 7; ++k)
            HIPCHK(hipMemcpyAsync(c->d_fields + (long)fl[k] * c->ld, f[k].data(), sizeof(double) * num_atoms,
                                  hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int sx_set_atom_values(sx_ctx *c, int field, int first1, int count, const double *values)
{
    return guard(c, [&] {
        need_system(c);
        if (field < 0 || field >= SX_NFIELDS || count < 0 || (count > 0 && !values))
            throw sx_error{SX_ERR_ARG, "bad field values"};
        if (first1 < 1 || first1 - 1 + count > c->n)
            throw sx_error{SX_ERR_ARG, "Error: Atom index " + std::to_string(first1) + " does not exist in system of " +
                                           std::to_string(c->n) + " atoms."};
        if (count == 0) return;
        HIPCHK(hipMemcpyAsync(c->d_fields + (long)field * c->ld + (first1 - 1), values, sizeof(double) * count,
                              hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int sx_get_atom_values(sx_ctx *c, int field, int first1, int count, double *values)
{
    return guard(c, [&] {
        need_system(c);
        if (field < 0 || field >= SX_NFIELDS || count < 0 || (count > 0 && !values))
            throw sx_error{SX_ERR_ARG, "bad field values"};
        if (first1 < 1 || first1 - 1 + count > c->n)
            throw sx_error{SX_ERR_ARG, "Error: Atom index " + std::to_string(first1) + " does not exist in system of " +
                                           std::to_string(c->n) + " atoms."};
        if (count == 0) return;
        HIPCHK(hipMemcpyAsync(values, c->d_fields + (long)field * c->ld + (first1 - 1), sizeof(double) * count,
                              hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int sx_set_atom_types(sx_ctx *c, const int *rtype)
{
    return guard(c, [&] {
        need_system(c);
        if (!rtype) throw sx_error{SX_ERR_ARG, "null types"};
        c->h_rtype.assign(rtype, rtype + c->n);
    });
}

int sx_get_atom_types(sx_ctx *c, int *rtype)
{
    return guard(c, [&] {
        need_system(c);
        if (!rtype) throw sx_error{SX_ERR_ARG, "null types"};
        std::copy(c->h_rtype.begin(), c->h_rtype.end(), rtype);
    });
}

int sx_get_box(sx_ctx *c, double box[3])
{
    return guard(c, [&] { box[0] = c->box[0]; box[1] = c->box[1]; box[2] = c->box[2]; });
}

int sx_upload(sx_ctx *c, int field, const double *host, int n)
{
    return guard(c, [&] {
        need_system(c);
        if (field < 0 || field >= SX_NFIELDS || n != c->n || !host) throw sx_error{SX_ERR_ARG, "bad upload"};
        HIPCHK(hipMemcpyAsync(c->d_fields + field * c->ld, host, sizeof(double) * n, hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int sx_download(sx_ctx *c, int field, double *host, int n)
{
    return guard(c, [&] {
        need_system(c);
        if (field < 0 || field >= SX_NFIELDS || n != c->n || !host) throw sx_error{SX_ERR_ARG, "bad download"};
        HIPCHK(hipMemcpyAsync(host, c->d_fields + field * c->ld, sizeof(double) * n, hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int sx_field_ptr(sx_ctx *c, int field, double **p)
{
    return guard(c, [&] {
        need_system(c);
        if (field < 0 || field >= SX_NFIELDS || !p) throw sx_error{SX_ERR_ARG, "bad field"};
        *p = c->d_fields + field * c->ld;
    });
}

int sx_set_scalar(sx_ctx *c, int id, double v)
{
    return guard(c, [&] {
        if (id < 0 || id >= SX_NSCALARS) throw sx_error{SX_ERR_ARG, "bad scalar id"};
        sx_sync_scalars_to_host(c);
        // captured Langevin steps hold T and its noise counter int(1000 T)
        // as launch constants: a new target temperature re-captures them
        if (id == SX_TARGET_TEMP && (!c->have_target_temp || c->h_scal[id] != v)) c->have_graph = false;
        c->h_scal[id] = v;
        if (id == SX_TARGET_TEMP) c->have_target_temp = true;
        sx_sync_scalars_to_device(c);
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int sx_get_scalar(sx_ctx *c, int id, double *v)
{
    return guard(c, [&] {
        if (id < 0 || id >= SX_NSCALARS || !v) throw sx_error{SX_ERR_ARG, "bad scalar id"};
        sx_check_device_error(c);
        sx_sync_scalars_to_host(c);
        *v = c->h_scal[id];
    });
}

int sx_initialize_system(sx_ctx *c, int remove_com)
{
    return guard(c, [&] {
        need_system(c);
        if (c->box[0] <= 0 || c->box[1] <= 0 || c->box[2] <= 0)
            throw sx_error{SX_ERR_STATE, "Error: Box dimensions not set. Set box dimensions before initializing system."};
        hipLaunchKernelGGL(k_init_defaults, dim3(sx_blocks(c->n, 256)), dim3(256), 0, c->stream, c->dstate());
        for (int &t : c->h_rtype)   // system.f90:590-593
            if (t < 0) t = 0;
        sx_sync_scalars_to_host(c);
        int dof = 6 * c->n - c->nc;
        if (remove_com) dof -= 3;
        c->h_scal[SX_DOF] = dof;
        if (c->nrigid > 1) {
            c->h_scal[SX_TARGET_TEMP] = c->h_scal[SX_TARGET_TEMP] * c->nrigid;
            c->have_graph = false;
        }
        sx_sync_scalars_to_device(c);
        sx_init_forces_impl(c);
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int sx_set_velocity_to_temperature(sx_ctx *c, double rtemp, int seed, int remove_com)
{
    return guard(c, [&] {
        need_system(c);
        int n = c->n;
        int nb = sx_blocks(n, 256);
        int temp_int = (int)(rtemp * 1000);
        DState s = c->dstate();
        hipLaunchKernelGGL(k_vel_normals, dim3(nb), dim3(256), 0, c->stream, s, temp_int, seed);
        int n_dof = 6 * n - c->nc;
        sx_ensure_part(c, nb, 3);
        if (remove_com) {
            hipLaunchKernelGGL(k_vsum_part, dim3(nb), dim3(256), 0, c->stream, s, c->d_part);
            hipLaunchKernelGGL(k_vel_com_sub, dim3(nb), dim3(256), 0, c->stream, s, c->d_part, nb);
            n_dof -= 3;
        }
        hipLaunchKernelGGL(k_vsq_part, dim3(nb), dim3(256), 0, c->stream, s, c->d_part);
        double numer = (double)n_dof * rtemp * (double)c->nrigid;
        hipLaunchKernelGGL(k_vel_scale, dim3(nb), dim3(256), 0, c->stream, s, c->d_part, nb, numer);
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int sx_initialize_forces(sx_ctx *c)
{
    return guard(c, [&] { need_system(c); sx_init_forces_impl(c); });
}

int sx_calculate_thermo(sx_ctx *c)
{
    return guard(c, [&] { need_system(c); sx_thermo_impl(c); });
}

int sx_set_constraints(sx_ctx *c, int nc, const int *bbi1, const int *bbj1, const double *len);

void sx_philox4x32_10(const uint32_t ctr[4], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = 0x6D2B79F5u, k1 = 0x1B873593u;
    for (int r = 1; r <= 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ (c1 ^ k0);
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ (c3 ^ k1);
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
        if (r < 10) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

int sx_philox_normals(sx_ctx *c, int n, int c2, int c3, int c4, double *out)
{
    return guard(c, [&] {
        if (n <= 0 || !out) throw sx_error{SX_ERR_ARG, "bad philox request"};
        double *d = nullptr;
        HIPCHK(hipMalloc(&d, sizeof(double) * 3 * n));
        hipLaunchKernelGGL(k_philox_normals, dim3(sx_blocks(n, 256)), dim3(256), 0, c->stream, n, c2, c3, c4, d);
        HIPCHK(hipMemcpyAsync(out, d, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(d));
    });
}

// sx_philox_normals with the device's scratch array as the staging buffer:
// no allocation.  The kernel writes all 3 n values it stages, so nothing is
// cleared first.  Calls on one GPU share its scratch: overlapping calls race.
int sx_philox_normals_noalloc(sx_ctx *c, int n, int c2, int c3, int c4, double *out)
{
    return guard(c, [&] {
        if (n <= 0 || !out) throw sx_error{SX_ERR_ARG, "bad philox request"};
        if (3L * n > SX_SCRATCH_DOUBLES)
            throw sx_error{SX_ERR_CAPACITY, "at most " + std::to_string(SX_SCRATCH_DOUBLES / 3) +
                                                " normal triples per call without allocation"};
        double *d = scratch_address(c->device);
        hipLaunchKernelGGL(k_philox_normals, dim3(sx_blocks(n, 256)), dim3(256), 0, c->stream, n, c2, c3, c4, d);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, d, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

}  // extern "C"

