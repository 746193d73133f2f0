// Device neighbour list: cell-binned build of the reference's Verlet list
// (module neighbour_list, system.f90:885-1053).
//
// The pair set is the reference's: a pair (a < b) is included iff the FP64
// value ((d0*d0 + d1*d1) + d2*d2), d = mimg(r_a - r_b), is < (rc+skin)^2,
// evaluated in the reference's order for the canonical orientation a < b.
// Cells (edge >= rc+skin, >= 3 per dimension) only choose which pairs are
// tested.  Each row holds every neighbour j != i sorted ascending ("full"
// list); the reference's half list (j > i) is the tail of each row.
#include "sx_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#define SORT_CAP 2048

// canonical pair test: identical FP64 expression to build_nlist (system.f90:980-989)
__device__ __forceinline__ bool in_list(const DState &s, int i, int j, double rsq_cut)
{
    int a = i < j ? i : j, b = i < j ? j : i;
    double d0 = s.f[SX_RX][a] - s.f[SX_RX][b];
    double d1 = s.f[SX_RY][a] - s.f[SX_RY][b];
    double d2 = s.f[SX_RZ][a] - s.f[SX_RZ][b];
    if (s.periodic) {
        d0 = sx_mimg(d0, s.box[0]);
        d1 = sx_mimg(d1, s.box[1]);
        d2 = sx_mimg(d2, s.box[2]);
    }
    double rsq = d0 * d0 + d1 * d1 + d2 * d2;
    return rsq < rsq_cut;
}

struct CellGrid {
    int nc[3];
    double w[3];     // cell edge
    double inv[3];   // 1 / box
    int fast;        // (rc+skin) < box/2: reciprocal min-image gives the reference's value
};

// the same test from packed positions (pa of particle i, pb of particle j)
__device__ __forceinline__ bool in_list_pos(double4 pi, double4 pj, int i, int j, const DState &s, const CellGrid &g,
                                            double rsq_cut)
{
    double4 pa = i < j ? pi : pj, pb = i < j ? pj : pi;
    double d0 = pa.x - pb.x, d1 = pa.y - pb.y, d2 = pa.z - pb.z;
    if (s.periodic) {
        if (g.fast) {
            d0 = sx_mimg_fast(d0, s.box[0], g.inv[0]);
            d1 = sx_mimg_fast(d1, s.box[1], g.inv[1]);
            d2 = sx_mimg_fast(d2, s.box[2], g.inv[2]);
        } else {
            d0 = sx_mimg(d0, s.box[0]);
            d1 = sx_mimg(d1, s.box[1]);
            d2 = sx_mimg(d2, s.box[2]);
        }
    }
    double rsq = d0 * d0 + d1 * d1 + d2 * d2;
    return rsq < rsq_cut;
}

__device__ __forceinline__ int cell_coord(double x, double L, int nc, double w)
{
    double xw = x - floor(x / L) * L;   // [0, L) for binning only
    int c = (int)(xw / w);
    return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

// bins the particles and records the positions the list is built from
// (x_last of build_nlist, system.f90:1001); count[7] counts the builds
// done inside graph replays
__device__ __forceinline__ void cell_assign_one(const DState &s, const CellGrid &g, int i, int *cell_of,
                                                int *cell_count, double *xl, double *yl, double *zl)
{
    double x = s.f[SX_RX][i], y = s.f[SX_RY][i], z = s.f[SX_RZ][i];
    xl[i] = x; yl[i] = y; zl[i] = z;
    int cx = cell_coord(x, s.box[0], g.nc[0], g.w[0]);
    int cy = cell_coord(y, s.box[1], g.nc[1], g.w[1]);
    int cz = cell_coord(z, s.box[2], g.nc[2], g.w[2]);
    int c = (cx * g.nc[1] + cy) * g.nc[2] + cz;
    cell_of[i] = c;
    atomicAdd(&cell_count[c], 1);
}

__global__ void k_cell_assign(DState s, CellGrid g, int *cell_of, int *cell_count, double *xl, double *yl,
                              double *zl, const int *flag, long *count)
{
    if (flag && *flag == 0) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && count) count[7] += 1;
    if (i >= s.n) return;
    cell_assign_one(s, g, i, cell_of, cell_count, xl, yl, zl);
}

// exclusive scan of n ints by one work-group into out[0..n-1]; returns the
// total (the same in every thread).  Tiles of 4096: 4 consecutive elements
// per thread, wavefront shuffle scans, one LDS pass over the 16 wavefront
// totals.
__device__ long scan_block(const int *in, int *out, int n, int *wsum)
{
    int t = threadIdx.x, lane = t & 63, w = t >> 6;
    long carry = 0;
    for (long base = 0; base < n; base += 4096) {
        long i0 = base + 4L * t;
        int v[4];
        for (int k = 0; k < 4; ++k) v[k] = (i0 + k < n) ? in[i0 + k] : 0;
        int local = v[0] + v[1] + v[2] + v[3];
        int x = local;
        for (int off = 1; off < 64; off <<= 1) {
            int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        if (w == 0) {
            int ws = lane < 16 ? wsum[lane] : 0;
            for (int off = 1; off < 16; off <<= 1) {
                int y = __shfl_up(ws, off, 64);
                if (lane >= off) ws += y;
            }
            if (lane < 16) wsum[lane] = ws;
        }
        __syncthreads();
        long excl = carry + (w > 0 ? wsum[w - 1] : 0) + (x - local);
        for (int k = 0; k < 4; ++k) {
            if (i0 + k < n) out[i0 + k] = (int)excl;
            excl += v[k];
        }
        long tile = wsum[15];
        __syncthreads();
        carry += tile;
    }
    return carry;
}

// exclusive scan of n ints by one work-group; out has n+1 entries; optional
// conditional flag (skip when *flag == 0).  With cap > 0 a total beyond cap
// leaves an empty list instead: out[0..n] (and zero_also[0..n-1]) all zero,
// *total the count that did not fit (for the host to grow the list) and the
// capacity error raised, so that no later kernel indexes past the list.
// in2 (may be NULL): a second array of n ints scanned the same way into
// out2[0..n] after the first (and after its zeroing), its total in *total2.
// the scan of k_scan below by one workgroup of 1024 threads
__device__ void scan_body(const int *in, int *out, int n, int *total, long cap, int *err, int *zero_also,
                          const int *in2, int *out2, int *total2, int *wsum)
{
    const int t = threadIdx.x;
    const long carry = scan_block(in, out, n, wsum);
    // carry is the same in every thread, so the branch is uniform
    if (cap > 0 && carry > cap) {
        for (long i = t; i <= n; i += blockDim.x) {
            out[i] = 0;
            if (zero_also && i < n) zero_also[i] = 0;
        }
        if (t == 0) {
            if (total) *total = carry > 0x7fffffffL ? 0x7fffffff : (int)carry;
            atomicCAS(err, 0, DERR_NLIST_CAP);
        }
        __syncthreads();   // the zeros are what the second scan reads
    } else if (t == 0) {
        out[n] = (int)carry;
        if (total) *total = (int)carry;
    }
    if (!in2) return;
    const long carry2 = scan_block(in2, out2, n, wsum);
    if (t == 0) {
        out2[n] = (int)carry2;
        if (total2) *total2 = (int)carry2;
    }
}

__global__ void __launch_bounds__(1024) k_scan(const int *in, int *out, int n, const int *flag, int *total,
                                               long cap, int *err, int *zero_also, const int *in2, int *out2,
                                               int *total2)
{
    if (flag && *flag == 0) return;
    __shared__ int wsum[16];
    scan_body(in, out, n, total, cap, err, zero_also, in2, out2, total2, wsum);
}

// places each particle in its cell's slot range by counting the cell's
// count back down, which leaves every count at zero for the next build
__device__ __forceinline__ void cell_scatter_one(const DState &s, int i, const int *cell_of, const int *cell_start,
                                                 int *cell_count, int *items, double4 *cell_pos)
{
    int c = cell_of[i];
    int p = cell_start[c + 1] - atomicSub(&cell_count[c], 1);
    items[p] = i;
    cell_pos[p] = make_double4(s.f[SX_RX][i], s.f[SX_RY][i], s.f[SX_RZ][i], 0.0);
}

__global__ void k_cell_scatter(DState s, const int *cell_of, const int *cell_start, int *cell_count, int *items,
                               double4 *cell_pos, const int *flag)
{
    if (flag && *flag == 0) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    cell_scatter_one(s, i, cell_of, cell_start, cell_count, items, cell_pos);
}

// count (pass 0) or fill (pass 1) the full rows using the cell grid
template <int PASS>
__global__ void k_rows_cells(DState s, CellGrid g, const int *cell_of, const int *cell_start, const int *items,
                             const double4 *cell_pos, double rsq_cut, int *row_len, int *lower, int *half_len,
                             const int *row_start, int *nbr, const int *flag, long cap, int *err)
{
    if (flag && *flag == 0) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    int c = cell_of[i];
    double4 pi = make_double4(s.f[SX_RX][i], s.f[SX_RY][i], s.f[SX_RZ][i], 0.0);
    int cz = c % g.nc[2], cy = (c / g.nc[2]) % g.nc[1], cx = c / (g.nc[2] * g.nc[1]);
    int cnt = 0, low = 0;
    long base = PASS ? row_start[i] : 0;
    if (PASS && (long)row_start[i + 1] > cap) {
        atomicCAS(err, 0, DERR_NLIST_CAP);
        return;
    }
    // a slot that does not match the counted length (the list was emptied
    // because it did not fit) is not filled
    if (PASS && row_start[i + 1] - row_start[i] != row_len[i]) return;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                int x = (cx + dx + g.nc[0]) % g.nc[0];
                int y = (cy + dy + g.nc[1]) % g.nc[1];
                int z = (cz + dz + g.nc[2]) % g.nc[2];
                int cc = (x * g.nc[1] + y) * g.nc[2] + z;
                for (int q = cell_start[cc]; q < cell_start[cc + 1]; ++q) {
                    int j = items[q];
                    if (j == i) continue;
                    if (in_list_pos(pi, cell_pos[q], i, j, s, g, rsq_cut)) {
                        if (PASS) nbr[base + cnt] = j;
                        ++cnt;
                        if (j < i) ++low;
                    }
                }
            }
    if (!PASS) {
        row_len[i] = cnt;
        lower[i] = low;
        half_len[i] = cnt - low;
    }
}

// The same rows with LPR lanes per particle.  For each (dx, dy) of the 3x3
// column of neighbour cells, the three z cells are consecutive in the
// cell-ordered item array (two ranges where z wraps), so a row is 9-18
// contiguous candidate ranges; the group's lanes test LPR candidates at a
// time and place hits with a group-masked ballot and prefix count.  Every loop
// bound is uniform within a group, so the group's ballots stay in step.
__device__ __forceinline__ unsigned long long group_mask(int LPR)
{
    const int l64 = threadIdx.x & 63;
    return LPR == 64 ? ~0ull : (((1ull << LPR) - 1ull) << (l64 - l64 % LPR));
}

// the row of the particle in cell-order slot `slot`, by the LPR lanes of a
// group (all of them with the same slot)
template <int PASS, int LPR>
__device__ __forceinline__ void rows_cells_group(const DState &s, const CellGrid &g, const int *cell_of,
                                                 const int *cell_start, const int *items, const double4 *cell_pos,
                                                 double rsq_cut, int *row_len, int *lower, int *half_len,
                                                 const int *row_start, int *nbr, long cap, int *err, int slot)
{
    const int lane = threadIdx.x % LPR;
    const int i = items[slot];
    const unsigned long long gm = group_mask(LPR);
    const unsigned long long below = gm & ((1ull << (threadIdx.x & 63)) - 1ull);
    const int c = cell_of[i];
    const double4 pi = cell_pos[slot];
    const int nx = g.nc[0], ny = g.nc[1], nz = g.nc[2];
    const int cz = c % nz, cy = (c / nz) % ny, cx = c / (nz * ny);
    int cnt = 0, low = 0;
    long base = 0;
    if (PASS) {
        base = row_start[i];
        if ((long)row_start[i + 1] > cap) {
            if (lane == 0) atomicCAS(err, 0, DERR_NLIST_CAP);
            return;
        }
        // a slot that does not match the counted length (the list was
        // emptied because it did not fit) is not filled; uniform per group
        if (row_start[i + 1] - row_start[i] != row_len[i]) return;
    }
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) {
            const int x = (cx + dx + nx) % nx, y = (cy + dy + ny) % ny;
            const int col = (x * ny + y) * nz;
            // z cells cz-1..cz+1 as up to two consecutive runs
            int r0[2], r1[2], nr = 0;
            const int zlo = cz - 1 < 0 ? 0 : cz - 1, zhi = cz + 1 > nz - 1 ? nz - 1 : cz + 1;
            r0[nr] = cell_start[col + zlo]; r1[nr] = cell_start[col + zhi + 1]; ++nr;
            if (cz - 1 < 0) { r0[nr] = cell_start[col + nz - 1]; r1[nr] = cell_start[col + nz]; ++nr; }
            else if (cz + 1 > nz - 1) { r0[nr] = cell_start[col]; r1[nr] = cell_start[col + 1]; ++nr; }
            for (int rr = 0; rr < nr; ++rr)
                for (int q0 = r0[rr]; q0 < r1[rr]; q0 += LPR) {
                    const int q = q0 + lane;
                    int j = -1;
                    bool hit = false;
                    if (q < r1[rr]) {
                        j = items[q];
                        hit = j != i && in_list_pos(pi, cell_pos[q], i, j, s, g, rsq_cut);
                    }
                    const unsigned long long m = __ballot(hit) & gm;
                    if (PASS && hit) nbr[base + cnt + __popcll(m & below)] = j;
                    cnt += __popcll(m);
                    if (!PASS) low += __popcll(__ballot(hit && j < i) & gm);
                }
        }
    if (!PASS && lane == 0) {
        row_len[i] = cnt;
        lower[i] = low;
        half_len[i] = cnt - low;
    }
}

template <int PASS, int LPR>
__global__ void __launch_bounds__(256) k_rows_cells_g(DState s, CellGrid g, const int *cell_of, const int *cell_start,
                                                      const int *items, const double4 *cell_pos, double rsq_cut,
                                                      int *row_len, int *lower, int *half_len, const int *row_start,
                                                      int *nbr, const int *flag, long cap, int *err)
{
    if (flag && *flag == 0) return;
    // rows are taken in cell order, so neighbouring groups read the same cells
    const int slot = blockIdx.x * (256 / LPR) + threadIdx.x / LPR;
    if (slot >= s.n) return;   // whole groups leave together
    rows_cells_group<PASS, LPR>(s, g, cell_of, cell_start, items, cell_pos, rsq_cut, row_len, lower, half_len,
                                row_start, nbr, cap, err, slot);
}

// half-pair index with LPR lanes per row (entries of a row are independent)
// row i's entries by lane `lane` of the row's L lanes
__device__ __forceinline__ void half_index_row(int n, int i, int lane, int L, const int *row_start, const int *lower,
                                               const int *half_start, const int *nbr, int *pairk, int *half_i,
                                               int *half_j, int *fwdpos, int *revpos, long cap_half, int *err)
{
    if ((long)half_start[n] > cap_half) {
        if (lane == 0) atomicCAS(err, 0, DERR_NLIST_CAP);
        return;
    }
    const int b = row_start[i], e = row_start[i + 1], lo = lower[i];
    for (int p = b + lane; p < e; p += L) {
        const int j = nbr[p];
        if (j > i) {
            const int k = half_start[i] + (p - b - lo);
            pairk[p] = k;
            half_i[k] = i;
            half_j[k] = j;
            fwdpos[k] = p;
        } else {
            // position of i in row j (sorted): binary search
            int lb = row_start[j], ub = row_start[j + 1];
            while (lb < ub) {
                const int m = (lb + ub) >> 1;
                if (nbr[m] < i) lb = m + 1; else ub = m;
            }
            const int k = half_start[j] + (lb - row_start[j] - lower[j]);
            pairk[p] = k;
            revpos[k] = p;
        }
    }
}

template <int LPR>
__global__ void __launch_bounds__(256) k_half_index_g(int n, const int *row_start, const int *lower,
                                                      const int *half_start, const int *nbr, int *pairk, int *half_i,
                                                      int *half_j, int *fwdpos, int *revpos, const int *flag,
                                                      long cap_half, int *err)
{
    if (flag && *flag == 0) return;
    const int i = blockIdx.x * (256 / LPR) + threadIdx.x / LPR;
    if (i >= n) return;
    half_index_row(n, i, threadIdx.x % LPR, LPR, row_start, lower, half_start, nbr, pairk, half_i, half_j, fwdpos,
                   revpos, cap_half, err);
}

// brute force rows (small or non-periodic boxes)
template <int PASS>
__global__ void k_rows_brute(DState s, double rsq_cut, int *row_len, int *lower, int *half_len, const int *row_start,
                             int *nbr, const int *flag, long cap, int *err)
{
    if (flag && *flag == 0) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    int cnt = 0, low = 0;
    long base = PASS ? row_start[i] : 0;
    if (PASS && (long)row_start[i + 1] > cap) {
        atomicCAS(err, 0, DERR_NLIST_CAP);
        return;
    }
    // a slot that does not match the counted length (the list was emptied
    // because it did not fit) is not filled
    if (PASS && row_start[i + 1] - row_start[i] != row_len[i]) return;
    for (int j = 0; j < s.n; ++j) {
        if (j == i) continue;
        if (in_list(s, i, j, rsq_cut)) {
            if (PASS) nbr[base + cnt] = j;
            ++cnt;
            if (j < i) ++low;
        }
    }
    if (!PASS) {
        row_len[i] = cnt;
        lower[i] = low;
        half_len[i] = cnt - low;
    }
}

// one wavefront per row: the row goes to LDS, each lane takes up to 8 of its
// entries (as few as the row's length allows) and ranks them against the whole row (the entries of a row are
// distinct, so rank = number of smaller entries is the sorted position), then
// writes them back in ascending j
#define SORT_PER_LANE 8

// PER entries of the row per lane from q0 on, ranked against the whole row
template <int PER>
__device__ __forceinline__ void rank_row_entries(const int *sh, int len, int len4, int lane, int q0, int *row)
{
    int mine[PER], rank[PER];
#pragma unroll
    for (int t = 0; t < PER; ++t) {
        const int q = q0 + t * SX_WAVE + lane;
        mine[t] = q < len ? sh[q] : 0x7fffffff;
        rank[t] = 0;
    }
    const int4 *sh4 = (const int4 *)sh;
    for (int m = 0; m < len4 / 4; ++m) {
        const int4 v = sh4[m];   // the same address in every lane: a broadcast read
#pragma unroll
        for (int t = 0; t < PER; ++t)
            rank[t] += (v.x < mine[t]) + (v.y < mine[t]) + (v.z < mine[t]) + (v.w < mine[t]);
    }
#pragma unroll
    for (int t = 0; t < PER; ++t) {
        const int q = q0 + t * SX_WAVE + lane;
        if (q < len) row[rank[t]] = mine[t];
    }
}

// row i sorted in place by one wavefront through its LDS buffer sh
__device__ __forceinline__ void sort_row(int i, const int *row_start, int *nbr, int *sh, int lane, int *err)
{
    {
        const int b = row_start[i], len = row_start[i + 1] - b;
        if (len <= 1) return;
        if (len > SORT_CAP) {
            if (lane == 0) atomicCAS(err, 0, DERR_ROW_CAP);
            return;
        }
        const int len4 = (len + 3) & ~3;
        for (int q = lane; q < len4; q += SX_WAVE) sh[q] = q < len ? nbr[b + q] : 0x7fffffff;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // as many entries per lane as the row needs (up to SORT_PER_LANE per pass)
        const int per = min(SORT_PER_LANE, (len + SX_WAVE - 1) / SX_WAVE);
        for (int q0 = 0; q0 < len; q0 += SX_WAVE * per) {
            switch (per) {
            case 1: rank_row_entries<1>(sh, len, len4, lane, q0, nbr + b); break;
            case 2: rank_row_entries<2>(sh, len, len4, lane, q0, nbr + b); break;
            case 3: rank_row_entries<3>(sh, len, len4, lane, q0, nbr + b); break;
            case 4: rank_row_entries<4>(sh, len, len4, lane, q0, nbr + b); break;
            case 5: rank_row_entries<5>(sh, len, len4, lane, q0, nbr + b); break;
            case 6: rank_row_entries<6>(sh, len, len4, lane, q0, nbr + b); break;
            case 7: rank_row_entries<7>(sh, len, len4, lane, q0, nbr + b); break;
            default: rank_row_entries<8>(sh, len, len4, lane, q0, nbr + b); break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

__global__ void __launch_bounds__(256) k_sort_rows(int n, const int *row_start, int *nbr, const int *flag, int *err)
{
    if (flag && *flag == 0) return;
    __shared__ __align__(16) int buf[4][SORT_CAP];
    const int w = threadIdx.x / SX_WAVE, lane = threadIdx.x % SX_WAVE;
    for (int i = blockIdx.x * 4 + w; i < n; i += gridDim.x * 4) sort_row(i, row_start, nbr, buf[w], lane, err);
}

// half-pair index of every full-list entry, and the half list (i, j) pairs
__global__ void k_half_index(int n, const int *row_start, const int *lower, const int *half_start, const int *nbr,
                             int *pairk, int *half_i, int *half_j, const int *flag, long cap_half, int *err)
{
    if (flag && *flag == 0) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int b = row_start[i], e = row_start[i + 1], lo = lower[i];
    if ((long)half_start[n] > cap_half) {
        atomicCAS(err, 0, DERR_NLIST_CAP);
        return;
    }
    for (int p = b; p < e; ++p) {
        int j = nbr[p];
        if (j > i) {
            int k = half_start[i] + (p - b - lo);
            pairk[p] = k;
            half_i[k] = i;
            half_j[k] = j;
        } else {
            // position of i in row j (sorted): binary search
            int lb = row_start[j], ub = row_start[j + 1];
            while (lb < ub) {
                int m = (lb + ub) >> 1;
                if (nbr[m] < i) lb = m + 1; else ub = m;
            }
            pairk[p] = half_start[j] + (lb - row_start[j] - lower[j]);
        }
    }
}

__global__ void k_save_last(DState s, double *xl, double *yl, double *zl, const int *flag, long *count)
{
    if (flag && *flag == 0) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && count) count[7] += 1;   // rebuilds done inside graph replays
    if (i >= s.n) return;
    xl[i] = s.f[SX_RX][i]; yl[i] = s.f[SX_RY][i]; zl[i] = s.f[SX_RZ][i];
}

// needs_rebuild (system.f90:1005-1039): max over i of |mimg(r - r_last)|^2
// needs_rebuild in one pass: per-block maxima, the last block to finish takes
// the maximum over blocks and writes the rebuild flag (max is order-free)
__global__ void __launch_bounds__(256) k_disp_max(DState s, const double *xl, const double *yl, const double *zl,
                                                  double *part, unsigned *ticket, double thresh, int rule, int *flag)
{
    __shared__ double sh[256];
    __shared__ bool last;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (i < s.n) {
        double dx = s.f[SX_RX][i] - xl[i], dy = s.f[SX_RY][i] - yl[i], dz = s.f[SX_RZ][i] - zl[i];
        if (s.periodic) {
            dx = sx_mimg(dx, s.box[0]); dy = sx_mimg(dy, s.box[1]); dz = sx_mimg(dz, s.box[2]);
        }
        m = dx * dx + dy * dy + dz * dz;
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
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
    for (int b = threadIdx.x; b < (int)gridDim.x; b += 256)
        v = fmax(v, __hip_atomic_load(&part[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double mx = sh[0];
        flag[0] = rule == SX_REBUILD_REFERENCE ? (4.0 * mx > thresh) : (mx > thresh);
        sx_ticket_reset(ticket);
    }
}

static CellGrid cell_grid(const sx_ctx *c);

// ---------------------------------------------------------------------------
// The conditional rebuild of a replayed step in one launch.  Between builds a
// step pays one launch that reads the rebuild flag and returns (eight flag-
// gated launches of the stage kernels above cost about 2 us each inside a
// replayed graph).  When the flag is set, the same stages run one after the
// other in LB_GRID resident workgroups of 1024 threads (one per CU), which
// meet at device-wide barriers between stages: bin, cell starts (workgroup
// 0), scatter, count rows, row and half-list starts (workgroup 0), fill rows,
// sort rows, half index.  The stage code is the kernels' own.
// ---------------------------------------------------------------------------
#define LB_THREADS 1024
#define LB_WAVES (LB_THREADS / SX_WAVE)
#define LB_GMAX 256          // most workgroups (four flags per lane of the polling wavefront)
#define LB_SPIN (1 << 16)    // polls of one barrier before DERR_LIST_SYNC

struct ListArgs {
    CellGrid g;
    int *cell_of, *cell_count, *cell_start, *cell_items;
    double4 *cell_pos;
    double *xl, *yl, *zl;
    long *count;
    int ncells;
    double rsq_cut;
    int *row_len, *lower, *half_len, *row_start, *nbr;
    long cap_full;
    int need_half, lpr_half;
    int *half_start, *pairk, *half_i, *half_j, *fwdpos, *revpos;
    long cap_half;
    int *total, *total_half;     // lengths of the full and half lists
    int *err;
    unsigned long long *sync;    // per workgroup: epoch * 16 + the last stage finished; [LB_GMAX]: the epoch
};

// Every workgroup has finished the stage that `target` names: each wavefront
// waits for its own stores, one L2 write-back per workgroup makes them
// visible to the other XCDs, then the flag; the first wavefront polls every
// flag (lane l: workgroups l, l + 64, l + 128, l + 192, loaded before any is
// tested), then one L2 invalidation before the workgroup reads what the
// others wrote.  false after LB_SPIN polls (DERR_LIST_SYNC raised), so that
// every wait ends.
__device__ bool lb_barrier(const ListArgs &a, unsigned long long target)
{
    __shared__ int ok_sh;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(a.sync + blockIdx.x, target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (threadIdx.x < SX_WAVE) {
        const int G = gridDim.x;
        bool done = false;
        for (int it = 0; it < LB_SPIN && !done; ++it) {
            unsigned long long f[LB_GMAX / SX_WAVE];
#pragma unroll
            for (int k = 0; k < LB_GMAX / SX_WAVE; ++k) {
                const int w = threadIdx.x + k * SX_WAVE;
                f[k] = w < G ? __hip_atomic_load(a.sync + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : target;
            }
            bool ok = true;
#pragma unroll
            for (int k = 0; k < LB_GMAX / SX_WAVE; ++k) ok = ok && f[k] >= target;
            done = __all(ok);
            if (!done) __builtin_amdgcn_s_sleep(1);
        }
        if (threadIdx.x == 0) {
            ok_sh = done;
            if (!done) atomicCAS(a.err, DERR_NONE, DERR_LIST_SYNC);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    return ok_sh != 0;
}

template <int LPR>
__global__ void __launch_bounds__(LB_THREADS) k_list_rebuild(DState s, ListArgs a, const int *flag)
{
    if (*flag == 0) return;
    __shared__ int wsum[16];
    __shared__ __align__(16) int sortbuf[LB_WAVES][SORT_CAP];
    const int n = s.n, G = gridDim.x, lane = threadIdx.x % SX_WAVE, w = threadIdx.x / SX_WAVE;
    const int tid = blockIdx.x * LB_THREADS + threadIdx.x, nth = G * LB_THREADS;
    const unsigned long long ep =
        __hip_atomic_load(a.sync + LB_GMAX, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) * 16ull;
    // bin the particles, recording the positions the list is built from
    if (tid == 0 && a.count) a.count[7] += 1;
    for (int i = tid; i < n; i += nth) cell_assign_one(s, a.g, i, a.cell_of, a.cell_count, a.xl, a.yl, a.zl);
    if (!lb_barrier(a, ep + 1)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0)   // every workgroup has started: the next build's epoch
        __hip_atomic_store(a.sync + LB_GMAX, ep / 16ull + 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // cell starts
    if (blockIdx.x == 0)
        scan_body(a.cell_count, a.cell_start, a.ncells, nullptr, 0L, nullptr, nullptr, nullptr, nullptr, nullptr,
                  wsum);
    if (!lb_barrier(a, ep + 2)) return;
    for (int i = tid; i < n; i += nth)
        cell_scatter_one(s, i, a.cell_of, a.cell_start, a.cell_count, a.cell_items, a.cell_pos);
    if (!lb_barrier(a, ep + 3)) return;
    // rows in cell order, LPR lanes per row
    for (int slot = tid / LPR; slot < n; slot += nth / LPR)
        rows_cells_group<0, LPR>(s, a.g, a.cell_of, a.cell_start, a.cell_items, a.cell_pos, a.rsq_cut, a.row_len,
                                 a.lower, a.half_len, a.row_start, a.nbr, a.cap_full, a.err, slot);
    if (!lb_barrier(a, ep + 4)) return;
    // row starts, and the half-list starts when the half list is kept
    if (blockIdx.x == 0)
        scan_body(a.row_len, a.row_start, n, a.total, a.cap_full, a.err, a.half_len,
                  a.need_half ? (const int *)a.half_len : nullptr, a.half_start, a.total_half, wsum);
    if (!lb_barrier(a, ep + 5)) return;
    for (int slot = tid / LPR; slot < n; slot += nth / LPR)
        rows_cells_group<1, LPR>(s, a.g, a.cell_of, a.cell_start, a.cell_items, a.cell_pos, a.rsq_cut, a.row_len,
                                 a.lower, a.half_len, a.row_start, a.nbr, a.cap_full, a.err, slot);
    if (!lb_barrier(a, ep + 6)) return;
    for (int i = blockIdx.x * LB_WAVES + w; i < n; i += G * LB_WAVES) sort_row(i, a.row_start, a.nbr, sortbuf[w], lane, a.err);
    if (!a.need_half) return;
    if (!lb_barrier(a, ep + 7)) return;
    const int L = a.lpr_half;
    for (int i = tid / L; i < n; i += nth / L)
        half_index_row(n, i, tid % L, L, a.row_start, a.lower, a.half_start, a.nbr, a.pairk, a.half_i, a.half_j,
                       a.fwdpos, a.revpos, a.cap_half, a.err);
}

template <int LPR>
static int lb_blocks_per_cu()
{
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_list_rebuild<LPR>, LB_THREADS, 0) != hipSuccess) return 0;
    return per;
}

// workgroups of the one-launch rebuild for this device and row width, 0 when
// a workgroup of it does not fit a CU
static int lb_grid_for(int lpr)
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    int per = 0;
    switch (lpr) {
    case 4: per = lb_blocks_per_cu<4>(); break;
    case 8: per = lb_blocks_per_cu<8>(); break;
    case 16: per = lb_blocks_per_cu<16>(); break;
    case 32: per = lb_blocks_per_cu<32>(); break;
    default: per = lb_blocks_per_cu<64>(); break;
    }
    return per >= 1 ? std::min(cus, LB_GMAX) : 0;
}

static void launch_list_rebuild(sx_ctx *c, const int *flag)
{
    NList &nl = c->nl;
    ListArgs a;
    a.g = cell_grid(c);
    a.cell_of = nl.cell_of; a.cell_count = nl.cell_count; a.cell_start = nl.cell_start; a.cell_items = nl.cell_items;
    a.cell_pos = (double4 *)nl.cell_pos;
    a.xl = nl.x_last; a.yl = nl.y_last; a.zl = nl.z_last;
    a.count = c->d_counts;
    a.ncells = nl.ncells;
    a.rsq_cut = nl.r_skin_sq;
    a.row_len = nl.row_len; a.lower = nl.lower; a.half_len = nl.half_len; a.row_start = nl.row_start; a.nbr = nl.nbr;
    a.cap_full = nl.cap_full;
    a.need_half = nl.need_half ? 1 : 0;
    a.lpr_half = nl.lpr_half;
    a.half_start = nl.half_start; a.pairk = nl.pairk; a.half_i = nl.half_i; a.half_j = nl.half_j;
    a.fwdpos = nl.fwdpos; a.revpos = nl.revpos;
    a.cap_half = nl.cap_half;
    a.total = c->d_flag + 1; a.total_half = c->d_flag + 2;
    a.err = c->d_err;
    a.sync = nl.lb_sync;
    DState s = c->dstate();
    dim3 grid(nl.lb_grid), blk(LB_THREADS);
    switch (nl.lpr_rows) {
    case 4: hipLaunchKernelGGL(k_list_rebuild<4>, grid, blk, 0, c->stream, s, a, flag); break;
    case 8: hipLaunchKernelGGL(k_list_rebuild<8>, grid, blk, 0, c->stream, s, a, flag); break;
    case 16: hipLaunchKernelGGL(k_list_rebuild<16>, grid, blk, 0, c->stream, s, a, flag); break;
    case 32: hipLaunchKernelGGL(k_list_rebuild<32>, grid, blk, 0, c->stream, s, a, flag); break;
    default: hipLaunchKernelGGL(k_list_rebuild<64>, grid, blk, 0, c->stream, s, a, flag); break;
    }
}

// after DERR_LIST_SYNC: a build whose workgroups gave up waiting left its
// flags mid-build
void sx_nlist_reset_sync(sx_ctx *c)
{
    NList &nl = c->nl;
    if (nl.lb_sync) HIPCHK(hipMemsetAsync(nl.lb_sync, 0, sizeof(unsigned long long) * (LB_GMAX + 1), c->stream));
}

// ---------------------------------------------------------------------------
static void launch_rows_g(sx_ctx *c, int pass, const CellGrid &g, const int *flag)
{
    NList &nl = c->nl;
    const int L = nl.lpr_rows;
    dim3 grid((c->n + 256 / L - 1) / (256 / L)), blk(256);
#define SX_ROWS_G(P, LL)                                                                                            \
    hipLaunchKernelGGL((k_rows_cells_g<P, LL>), grid, blk, 0, c->stream, c->dstate(), g, nl.cell_of, nl.cell_start, \
                       nl.cell_items, (const double4 *)nl.cell_pos, nl.r_skin_sq, nl.row_len, nl.lower, nl.half_len,  \
                       nl.row_start, nl.nbr, flag, nl.cap_full, c->d_err)
    if (pass == 0) {
        switch (L) {
        case 4: SX_ROWS_G(0, 4); break;
        case 8: SX_ROWS_G(0, 8); break;
        case 16: SX_ROWS_G(0, 16); break;
        case 32: SX_ROWS_G(0, 32); break;
        default: SX_ROWS_G(0, 64); break;
        }
    } else {
        switch (L) {
        case 4: SX_ROWS_G(1, 4); break;
        case 8: SX_ROWS_G(1, 8); break;
        case 16: SX_ROWS_G(1, 16); break;
        case 32: SX_ROWS_G(1, 32); break;
        default: SX_ROWS_G(1, 64); break;
        }
    }
#undef SX_ROWS_G
}

static void launch_half_index_g(sx_ctx *c, const int *flag)
{
    NList &nl = c->nl;
    const int L = nl.lpr_half;
    dim3 grid((c->n + 256 / L - 1) / (256 / L)), blk(256);
#define SX_HALF_G(LL)                                                                                               \
    hipLaunchKernelGGL((k_half_index_g<LL>), grid, blk, 0, c->stream, c->n, nl.row_start, nl.lower, nl.half_start, \
                       nl.nbr, nl.pairk, nl.half_i, nl.half_j, nl.fwdpos, nl.revpos, flag, nl.cap_half, c->d_err)
    switch (L) {
    case 4: SX_HALF_G(4); break;
    case 8: SX_HALF_G(8); break;
    case 16: SX_HALF_G(16); break;
    case 32: SX_HALF_G(32); break;
    default: SX_HALF_G(64); break;
    }
#undef SX_HALF_G
}

static int pow2_clamp(double x, int lo, int hi)
{
    int p = lo;
    while (p < hi && p < x) p <<= 1;
    return p;
}

static void nl_free(NList &nl)
{
    int *ip[] = {nl.cell_of, nl.cell_count, nl.cell_start, nl.cell_items, nl.row_len, nl.row_start, nl.lower,
                 nl.half_start, nl.half_len, nl.nbr, nl.pairk, nl.half_i, nl.half_j, nl.fwdpos, nl.revpos};
    for (int *p : ip)
        if (p) hipFree(p);
    double *dp[] = {nl.x_last, nl.y_last, nl.z_last, nl.cell_pos};
    for (double *p : dp)
        if (p) hipFree(p);
    nl.cell_of = nl.cell_count = nl.cell_start = nl.cell_items = nl.row_len = nl.row_start = nl.lower = nullptr;
    nl.half_start = nl.half_len = nl.nbr = nl.pairk = nl.half_i = nl.half_j = nl.fwdpos = nl.revpos = nullptr;
    nl.x_last = nl.y_last = nl.z_last = nl.cell_pos = nullptr;
    if (nl.lb_sync) hipFree(nl.lb_sync);
    nl.lb_sync = nullptr;
    nl.lb_grid = 0;
}

static void nl_alloc_lists(sx_ctx *c, long cap_full)
{
    NList &nl = c->nl;
    if (nl.nbr) hipFree(nl.nbr);
    if (nl.pairk) hipFree(nl.pairk);
    if (nl.half_i) hipFree(nl.half_i);
    if (nl.half_j) hipFree(nl.half_j);
    if (nl.fwdpos) hipFree(nl.fwdpos);
    if (nl.revpos) hipFree(nl.revpos);
    if (cap_full > 0x7fffffffL) throw sx_error{SX_ERR_CAPACITY, "neighbour list larger than 2^31 entries"};
    nl.cap_full = cap_full;
    nl.cap_half = cap_full / 2 + 1;
    HIPCHK(hipMalloc(&nl.nbr, sizeof(int) * nl.cap_full));
    HIPCHK(hipMalloc(&nl.pairk, sizeof(int) * nl.cap_full));
    HIPCHK(hipMalloc(&nl.half_i, sizeof(int) * nl.cap_half));
    HIPCHK(hipMalloc(&nl.half_j, sizeof(int) * nl.cap_half));
    HIPCHK(hipMalloc(&nl.fwdpos, sizeof(int) * nl.cap_half));
    HIPCHK(hipMalloc(&nl.revpos, sizeof(int) * nl.cap_half));
    c->have_graph = false;
}

// grows the list arrays to cap_full entries keeping their contents (the list
// stays valid, the rebuild decision is not disturbed)
static void nl_grow_keep(sx_ctx *c, long cap_full)
{
    NList &nl = c->nl;
    if (cap_full <= nl.cap_full) return;
    if (cap_full > 0x7fffffffL) throw sx_error{SX_ERR_CAPACITY, "neighbour list larger than 2^31 entries"};
    const long cap_half = cap_full / 2 + 1;
    int **full[] = {&nl.nbr, &nl.pairk};
    int **half[] = {&nl.half_i, &nl.half_j, &nl.fwdpos, &nl.revpos};
    for (int **a : full) {
        int *p = nullptr;
        HIPCHK(hipMalloc(&p, sizeof(int) * cap_full));
        HIPCHK(hipMemcpyAsync(p, *a, sizeof(int) * nl.cap_full, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(*a));
        *a = p;
    }
    for (int **a : half) {
        int *p = nullptr;
        HIPCHK(hipMalloc(&p, sizeof(int) * cap_half));
        HIPCHK(hipMemcpyAsync(p, *a, sizeof(int) * nl.cap_half, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(*a));
        *a = p;
    }
    nl.cap_full = cap_full;
    nl.cap_half = cap_half;
    c->have_graph = false;
}

// entries of the last build (d_flag[1]; a build that did not fit leaves the
// count it needed there)
static long nl_last_total(sx_ctx *c)
{
    int total = 0;
    HIPCHK(hipMemcpyAsync(&total, c->d_flag + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return total;
}

// Room for the list to grow by half before graph replays, whose rebuilds
// cannot allocate.
void sx_nlist_headroom(sx_ctx *c)
{
    NList &nl = c->nl;
    if (!nl.init) return;
    const long total = nl_last_total(c);
    if (3 * total / 2 + 1024 > nl.cap_full) nl_grow_keep(c, 2 * total + 2048);
}

// After graph replays: a rebuild that did not fit emptied the list (k_scan)
// and raised DERR_NLIST_CAP.  The list is then grown to the size it needed,
// rebuilt at the current positions, and the capacity error reported; any
// other device error is reported as it is.
void sx_nlist_check_after_replay(sx_ctx *c, bool headroom)
{
    int e = 0;
    HIPCHK(hipMemcpyAsync(&e, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    NList &nl = c->nl;
    if (e == DERR_NLIST_CAP && nl.init) {
        const long need = nl_last_total(c);
        const long had = nl.cap_full;
        HIPCHK(hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream));
        nl_alloc_lists(c, std::max(had, 3 * need / 2 + 1024));
        sx_nlist_build_impl(c, false);
        throw sx_error{SX_ERR_CAPACITY,
                       "neighbour list capacity exceeded during the run (" + std::to_string(need) + " entries, room for " +
                           std::to_string(had) + "): from that rebuild on the list was empty and the steps ran "
                           "without pair forces; the list has been grown and rebuilt at the current positions"};
    }
    sx_check_device_error(c);
    if (headroom) sx_nlist_headroom(c);
}

static void nl_setup(sx_ctx *c)
{
    NList &nl = c->nl;
    nl_free(nl);
    int n = c->n;
    double rskin = nl.cutoff + nl.skin;
    nl.r_skin_sq = rskin * rskin;
    nl.use_cells = c->periodic != 0;
    nl.ncells = 1;
    for (int d = 0; d < 3; ++d) {
        nl.ncell[d] = (int)std::floor(c->box[d] / rskin);
        if (nl.ncell[d] < 3) nl.use_cells = false;
    }
    if (nl.use_cells) {
        // bound the grid (very dilute systems)
        while ((long)nl.ncell[0] * nl.ncell[1] * nl.ncell[2] > 4L * n + 64) {
            for (int d = 0; d < 3; ++d) nl.ncell[d] = std::max(3, nl.ncell[d] / 2);
            if (nl.ncell[0] == 3 && nl.ncell[1] == 3 && nl.ncell[2] == 3) break;
        }
        nl.ncells = nl.ncell[0] * nl.ncell[1] * nl.ncell[2];
    }
    HIPCHK(hipMalloc(&nl.cell_of, sizeof(int) * n));
    HIPCHK(hipMalloc(&nl.cell_count, sizeof(int) * (nl.ncells + 1)));
    HIPCHK(hipMemsetAsync(nl.cell_count, 0, sizeof(int) * (nl.ncells + 1), c->stream));
    HIPCHK(hipMalloc(&nl.cell_start, sizeof(int) * (nl.ncells + 1)));
    HIPCHK(hipMalloc(&nl.cell_items, sizeof(int) * n));
    HIPCHK(hipMalloc(&nl.cell_pos, sizeof(double) * 4 * (size_t)n));
    HIPCHK(hipMalloc(&nl.row_len, sizeof(int) * (n + 1)));
    HIPCHK(hipMalloc(&nl.row_start, sizeof(int) * (n + 1)));
    HIPCHK(hipMalloc(&nl.lower, sizeof(int) * (n + 1)));
    HIPCHK(hipMalloc(&nl.half_start, sizeof(int) * (n + 1)));
    HIPCHK(hipMalloc(&nl.half_len, sizeof(int) * (n + 1)));
    HIPCHK(hipMalloc(&nl.x_last, sizeof(double) * n));
    HIPCHK(hipMalloc(&nl.y_last, sizeof(double) * n));
    HIPCHK(hipMalloc(&nl.z_last, sizeof(double) * n));
    // initial capacity: density estimate of the list sphere x 2, at least 32 per row
    double vol = c->box[0] * c->box[1] * c->box[2];
    double est = (double)n * (n / vol) * 4.18879 * rskin * rskin * rskin;
    long cap = std::max<long>((long)(2.0 * est) + 1024, 32L * n);
    if (!c->periodic) cap = std::max<long>(cap, std::min<long>((long)n * (n - 1), 64L * n));
    nl_alloc_lists(c, cap);
    // lanes per row: about the candidates of three cells for the row builder,
    // a quarter of the expected row for the half index
    double row_est = est / std::max(1, n);
    nl.lpr_rows = nl.use_cells ? pow2_clamp(3.0 * (double)n / nl.ncells, 4, 64) : 16;
    nl.lpr_half = pow2_clamp(row_est / 4.0, 4, 64);
    if (const char *e = getenv("SUCROSE_NLIST_LANES")) {   // rows,half override for tuning runs
        int a = 0, b = 0;
        if (sscanf(e, "%d,%d", &a, &b) == 2) {
            if (a >= 1) nl.lpr_rows = pow2_clamp(a, 4, 64);
            if (b >= 1) nl.lpr_half = pow2_clamp(b, 4, 64);
        }
    }
    // the one-launch conditional rebuild (cell grids only)
    if (nl.use_cells) {
        nl.lb_grid = lb_grid_for(nl.lpr_rows);
        if (nl.lb_grid > 0) {
            HIPCHK(hipMalloc(&nl.lb_sync, sizeof(unsigned long long) * (LB_GMAX + 1)));
            HIPCHK(hipMemsetAsync(nl.lb_sync, 0, sizeof(unsigned long long) * (LB_GMAX + 1), c->stream));
        }
    }
    nl.init = true;
}

// the reference's half list (half_start, half_i, half_j; d_flag[2] = its
// length) and pairk, from the sorted rows
static void half_index_launch(sx_ctx *c, const int *flag)
{
    NList &nl = c->nl;
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, c->stream, nl.half_len, nl.half_start, c->n, flag,
                       c->d_flag + 2, 0L, (int *)nullptr, (int *)nullptr, (const int *)nullptr, (int *)nullptr,
                       (int *)nullptr);
    launch_half_index_g(c, flag);
}

void sx_nlist_want_half(sx_ctx *c)
{
    NList &nl = c->nl;
    if (nl.need_half) return;
    nl.need_half = true;
    c->have_graph = false;
    if (!nl.init) return;
    half_index_launch(c, nullptr);
    int nh = 0;
    HIPCHK(hipMemcpyAsync(&nh, c->d_flag + 2, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    sx_check_device_error(c);
    nl.nnz_half = nh;
}

// Builds the list.  conditional: every kernel reads the device rebuild flag
// (graph replay); otherwise the build is unconditional and capacity grows as
// needed.
static CellGrid cell_grid(const sx_ctx *c)
{
    const NList &nl = c->nl;
    CellGrid g;
    g.fast = 1;
    for (int d = 0; d < 3; ++d) {
        g.nc[d] = nl.ncell[d];
        g.w[d] = c->box[d] / nl.ncell[d];
        g.inv[d] = 1.0 / c->box[d];
        if (!(nl.cutoff + nl.skin < 0.5 * c->box[d] * (1.0 - 1e-12))) g.fast = 0;
    }
    return g;
}

void sx_nlist_build_impl(sx_ctx *c, bool conditional)
{
    NList &nl = c->nl;
    int n = c->n;
    DState s = c->dstate();
    const int *flag = conditional ? c->d_flag : nullptr;
    int nb = sx_blocks(n, 256);
    if (conditional && nl.use_cells && nl.lb_grid > 0) {
        launch_list_rebuild(c, flag);
        return;
    }
    for (int attempt = 0; attempt < 4; ++attempt) {
        if (nl.use_cells) {
            CellGrid g = cell_grid(c);
            // the cell counts are zero between builds (k_cell_scatter counts
            // them back down)
            hipLaunchKernelGGL(k_cell_assign, dim3(nb), dim3(256), 0, c->stream, s, g, nl.cell_of, nl.cell_count,
                               nl.x_last, nl.y_last, nl.z_last, flag, conditional ? c->d_counts : (long *)nullptr);
            hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, c->stream, nl.cell_count, nl.cell_start, nl.ncells,
                               flag, (int *)nullptr, 0L, (int *)nullptr, (int *)nullptr, (const int *)nullptr,
                               (int *)nullptr, (int *)nullptr);
            hipLaunchKernelGGL(k_cell_scatter, dim3(nb), dim3(256), 0, c->stream, s, nl.cell_of, nl.cell_start,
                               nl.cell_count, nl.cell_items, (double4 *)nl.cell_pos, flag);
            launch_rows_g(c, 0, g, flag);
        } else {
            hipLaunchKernelGGL(k_rows_brute<0>, dim3(nb), dim3(256), 0, c->stream, s, nl.r_skin_sq, nl.row_len,
                               nl.lower, nl.half_len, nl.row_start, nl.nbr, flag, nl.cap_full, c->d_err);
        }
        // row starts, and in the same launch the half-list starts (d_flag[2]
        // = the half list's length) when the half list is kept
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, c->stream, nl.row_len, nl.row_start, n, flag,
                           c->d_flag + 1, nl.cap_full, c->d_err, nl.half_len,
                           nl.need_half ? (const int *)nl.half_len : (const int *)nullptr, nl.half_start,
                           c->d_flag + 2);
        if (nl.use_cells) {
            CellGrid g = cell_grid(c);
            launch_rows_g(c, 1, g, flag);
        } else {
            hipLaunchKernelGGL(k_rows_brute<1>, dim3(nb), dim3(256), 0, c->stream, s, nl.r_skin_sq, nl.row_len,
                               nl.lower, nl.half_len, nl.row_start, nl.nbr, flag, nl.cap_full, c->d_err);
        }
        if (!conditional) {
            int total = 0;
            HIPCHK(hipMemcpyAsync(&total, c->d_flag + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            if ((long)total > nl.cap_full) {
                HIPCHK(hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream));
                nl_alloc_lists(c, (long)(total * 1.5) + 1024);
                continue;
            }
        }
        break;
    }
    hipLaunchKernelGGL(k_sort_rows, dim3(std::min(sx_blocks(n, 4), 2048)), dim3(256), 0, c->stream, n, nl.row_start,
                       nl.nbr, flag, c->d_err);
    if (nl.need_half) launch_half_index_g(c, flag);   // half_start came with the row scan
    if (!nl.use_cells)
        hipLaunchKernelGGL(k_save_last, dim3(nb), dim3(256), 0, c->stream, s, nl.x_last, nl.y_last, nl.z_last, flag,
                           conditional ? c->d_counts : (long *)nullptr);
    if (!conditional) {
        int sizes[2];
        HIPCHK(hipMemcpyAsync(sizes, c->d_flag + 1, sizeof(int) * 2, hipMemcpyDeviceToHost, c->stream));
        sx_check_device_error(c);
        nl.nnz_full = sizes[0];
        nl.nnz_half = sizes[1];
        nl.rebuilds++;
    }
}

// computes the rebuild flag on the device (d_flag[0])
void sx_nlist_flag_impl(sx_ctx *c)
{
    NList &nl = c->nl;
    int nb = sx_blocks(c->n, 256);
    sx_ensure_part(c, nb, 1);
    double thresh = nl.rule == SX_REBUILD_REFERENCE ? nl.r_skin_sq : (nl.skin / 2.0) * (nl.skin / 2.0);
    hipLaunchKernelGGL(k_disp_max, dim3(nb), dim3(256), 0, c->stream, c->dstate(), nl.x_last, nl.y_last, nl.z_last,
                       c->d_part, c->d_ticket, thresh, nl.rule, c->d_flag);
}

static void need_nlist(sx_ctx *c)
{
    if (!c->nl.init) throw sx_error{SX_ERR_STATE, "neighbour list not initialised (sx_nlist_init)"};
}

// the list after a change of box (the barostat): the cell grid is derived
// again when its cells would no longer be at least rc+skin wide, then the
// list is built
void sx_nlist_rebox_impl(sx_ctx *c)
{
    NList &nl = c->nl;
    for (int d = 0; d < 3; ++d)
        if (nl.cutoff > c->box[d] / 2.0) throw sx_error{SX_ERR_CUTOFF, "A cutoff is too large for the box dimensions"};
    double rskin = nl.cutoff + nl.skin;
    bool same = nl.use_cells == (c->periodic != 0);
    for (int d = 0; d < 3 && same; ++d) {
        int want = (int)std::floor(c->box[d] / rskin);
        if (nl.use_cells ? (c->box[d] / nl.ncell[d] < rskin) : (want >= 3)) same = false;
    }
    if (!same) nl_setup(c);
    sx_nlist_build_impl(c, false);
    c->have_graph = false;
}

extern "C" {

int sx_nlist_init(sx_ctx *c, double skin, double cutoff)
{
    return guard(c, [&] {
        need_system(c);
        for (int d = 0; d < 3; ++d)
            if (cutoff > c->box[d] / 2.0) throw sx_error{SX_ERR_CUTOFF, "A cutoff is too large for the box dimensions"};
        c->nl.cutoff = cutoff;
        c->nl.skin = skin;
        nl_setup(c);
        sx_nlist_build_impl(c, false);
    });
}

int sx_nlist_build(sx_ctx *c)
{
    return guard(c, [&] { need_nlist(c); sx_nlist_build_impl(c, false); });
}

int sx_nlist_needs_rebuild(sx_ctx *c, int *out)
{
    return guard(c, [&] {
        need_nlist(c);
        sx_nlist_flag_impl(c);
        int f = 0;
        HIPCHK(hipMemcpyAsync(&f, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        *out = f;
    });
}

int sx_nlist_update(sx_ctx *c)
{
    return guard(c, [&] {
        need_nlist(c);
        sx_nlist_flag_impl(c);
        int f = 0;
        HIPCHK(hipMemcpyAsync(&f, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (f) sx_nlist_build_impl(c, false);
    });
}

int sx_nlist_set_rule(sx_ctx *c, int rule)
{
    return guard(c, [&] {
        if (rule != SX_REBUILD_REFERENCE && rule != SX_REBUILD_HALF_SKIN) throw sx_error{SX_ERR_ARG, "bad rule"};
        c->nl.rule = rule;
        c->have_graph = false;
    });
}

int sx_nlist_rebuild_count(sx_ctx *c, long *count)
{
    return guard(c, [&] {
        long dev[8];
        HIPCHK(hipMemcpyAsync(dev, c->d_counts, sizeof(long) * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        *count = c->nl.rebuilds + dev[7];
    });
}

int sx_nlist_export(sx_ctx *c, int *offset1, int *nlist1, long cap, long *nnz)
{
    return guard(c, [&] {
        need_nlist(c);
        sx_nlist_want_half(c);
        sx_check_device_error(c);
        int n = c->n;
        int sizes[2];
        HIPCHK(hipMemcpyAsync(sizes, c->d_flag + 1, sizeof(int) * 2, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        long nf = sizes[0], nh = sizes[1];
        *nnz = nh;
        std::vector<int> hs(n + 1), hi(nh), hj(nh);
        HIPCHK(hipMemcpyAsync(hs.data(), c->nl.half_start, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, c->stream));
        if (nh > 0) {
            HIPCHK(hipMemcpyAsync(hj.data(), c->nl.half_j, sizeof(int) * nh, hipMemcpyDeviceToHost, c->stream));
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        (void)nf;
        if (offset1) {
            for (int i = 0; i < n - 1; ++i) offset1[i] = hs[i] + 1;
            if (n >= 1) offset1[n - 1] = (int)nh + 1;
        }
        if (nlist1) {
            if (cap < nh) throw sx_error{SX_ERR_CAPACITY, "export buffer too small"};
            for (long k = 0; k < nh; ++k) nlist1[k] = hj[k] + 1;
        }
    });
}

}  // extern "C"
