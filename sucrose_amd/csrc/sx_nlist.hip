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

// half-pair index with LPR lanes per row: row i's upper entries (j > i, the
// tail of the sorted row) are half pairs half_start[i] onwards, in order
__device__ __forceinline__ void half_index_row(int n, int i, int lane, int L, const int *row_start, const int *lower,
                                               const int *half_start, const int *nbr, int *half_i, int *half_j,
                                               long cap_half, int *err)
{
    if ((long)half_start[n] > cap_half) {
        if (lane == 0) atomicCAS(err, 0, DERR_NLIST_CAP);
        return;
    }
    const int b = row_start[i] + lower[i], e = row_start[i + 1], h = half_start[i];
    for (int p = b + lane; p < e; p += L) {
        half_i[h + (p - b)] = i;
        half_j[h + (p - b)] = nbr[p];
    }
}

template <int LPR>
__global__ void __launch_bounds__(256) k_half_index_g(int n, const int *row_start, const int *lower,
                                                      const int *half_start, const int *nbr, int *half_i, int *half_j,
                                                      const int *flag, long cap_half, int *err)
{
    if (flag && *flag == 0) return;
    const int i = blockIdx.x * (256 / LPR) + threadIdx.x / LPR;
    if (i >= n) return;
    half_index_row(n, i, threadIdx.x % LPR, LPR, row_start, lower, half_start, nbr, half_i, half_j, cap_half, err);
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

// A row of at most 64 PER entries sorted in place by one wavefront with a
// bitonic network in registers: lane l holds entries l PER .. l PER + PER - 1
// (padded with INT_MAX), exchanges at a distance of PER or more go across
// lanes by shuffles, the others stay in the lane
template <int PER>
__device__ __forceinline__ void bitonic_row(int *row, int len, int lane)
{
    constexpr int N = SX_WAVE * PER;
    int v[PER];
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int q = lane * PER + r;
        v[r] = q < len ? row[q] : 0x7fffffff;
    }
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= PER) {
                const int lj = j / PER;
                const bool lower = (lane & lj) == 0;
#pragma unroll
                for (int r = 0; r < PER; ++r) {
                    const bool up = ((lane * PER + r) & k) == 0;
                    const int o = __shfl_xor(v[r], lj, SX_WAVE);
                    v[r] = lower == up ? min(v[r], o) : max(v[r], o);
                }
            } else {
#pragma unroll
                for (int r = 0; r < PER; ++r) {
                    const int p = r ^ j;
                    if (p > r) {
                        const bool up = ((lane * PER + r) & k) == 0;
                        const int a = v[r], b = v[p];
                        const bool sw = (a > b) == up;
                        v[r] = sw ? b : a;
                        v[p] = sw ? a : b;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int q = lane * PER + r;
        if (q < len) row[q] = v[r];
    }
}

// row i sorted in place by one wavefront: up to 512 entries in registers
// (bitonic_row), longer rows through its LDS buffer sh by ranks
__device__ __forceinline__ void sort_row(int i, const int *row_start, int *nbr, int *sh, int lane, int *err)
{
    {
        const int b = row_start[i], len = row_start[i + 1] - b;
        if (len <= 1) return;
        if (len <= 8 * SX_WAVE) {
            if (len <= SX_WAVE) bitonic_row<1>(nbr + b, len, lane);
            else if (len <= 2 * SX_WAVE) bitonic_row<2>(nbr + b, len, lane);
            else if (len <= 4 * SX_WAVE) bitonic_row<4>(nbr + b, len, lane);
            else bitonic_row<8>(nbr + b, len, lane);
            return;
        }
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
static bool lb_scratch(sx_ctx *c, int rc);
static void lb_bins(sx_ctx *c, int cap);
static int nl_max_row(sx_ctx *c);

// ---------------------------------------------------------------------------
// The conditional rebuild of a replayed step in one launch.  Between builds a
// step pays one launch that reads the rebuild flag and returns.  When the
// flag is set, one resident workgroup of LB_THREADS threads per CU runs the
// stages one after the other, meeting at LB_NBAR device-wide barriers:
//   1. bin: each particle takes a slot of its cell (fixed room lb_cap per
//      cell, an atomic count), recording the positions the list is built
//      from;
//   2. rows, four cells at a time per workgroup (cells handed out in
//      batches by a counter), each by a quarter of the wavefronts: the
//      candidates of the cell's 27 neighbour cells are staged in LDS once for
//      all of the cell's particles; a wavefront takes a particle, tests every
//      candidate with the reference's pair test, compacts the hits into LDS,
//      ranks them (the row sorted ascending) and writes the row to a scratch
//      area of fixed stride lb_rc;
//   3. the row lengths summed over each workgroup's contiguous range of
//      particles;
//   4. each workgroup's offset from the range sums before it, the row (and
//      half-list) starts of its range and a packed record per particle (row
//      start, half start, lower count, row length); a total beyond the
//      list's capacity, a cell beyond lb_cap or a row beyond lb_rc leaves an
//      empty list for the host to grow (as the stage kernels do);
//   5. the rows copied from the scratch into the list in particle order, and
//      the half pairs (i, j > i) at the places the row's record gives.
// The pair set and the order of every row are those of the stage kernels.
//
// Barrier: every workgroup stores its arrival word (agent scope); the first
// wavefront of workgroup 0 polls all of them, then stores LB_NREL release
// words, each in its own cache line, which the workgroups poll (a sixteenth
// of them each).  The bins, rows and row records (stages 1, 2 and 4) are
// written with plain stores and one L2 write-back per workgroup before its
// arrival; the range sums of stage 3 with agent-scope stores, written through
// to memory, so that barrier needs no write-back.  A workgroup invalidates
// its L2 and L1 once after each barrier before it reads.  Every wait gives up
// after LB_SPIN polls (DERR_LIST_SYNC, the build abandoned), so that every
// wait ends.
// ---------------------------------------------------------------------------
#define LB_THREADS 1024
#define LB_WAVES (LB_THREADS / SX_WAVE)
#define LB_NQ 4              // cells in flight per workgroup (a quarter of the wavefronts each)
#define LB_QW (LB_WAVES / LB_NQ)
#define LB_QT (LB_THREADS / LB_NQ)
#define LB_TMAX 512          // candidates staged in LDS per cell (more: read from the bins)
#define LB_RB 512            // longest row (LDS row buffer per wavefront)
#define LB_NBAR 4            // barriers per build
#define LB_NREL 16           // release words
#define LB_SPIN (1 << 20)    // polls of one barrier before DERR_LIST_SYNC
#define LB_ABORT 0xffffffffu
// sync words: [0] build epoch, the cell-batch counters of even and odd
// epochs at LB_CTR0 and LB_CTR0 + 32, release word r at LB_REL0 + 32 r,
// arrival of workgroup b at LB_ARR0 + b
#define LB_CTR0 32
#define LB_REL0 96
#define LB_ARR0 (LB_REL0 + 32 * LB_NREL)
#define LB_SYNC_WORDS (LB_ARR0 + 4096)

struct ListArgs {
    CellGrid g;
    int *cell_count;
    int *bin_items;              // ncells x cap
    double4 *bin_pos;
    int cap;
    double *xl, *yl, *zl;
    long *count;
    int ncells;
    double rsq_cut;
    int *row_len, *lower, *half_len, *row_start, *nbr;
    long cap_full;
    int need_half;
    int *half_start, *half_i, *half_j;
    int *total, *total_half, *max_row, *max_cell;   // d_flag[1], [2], [3], [5]
    int *rows;                   // scratch rows, stride rc
    int rc;
    int4 *meta;                  // per particle: row start, half start, lower count, row length
    int *bsum;                   // per workgroup: full and half row sums of its range
    int *err;
    unsigned *sync;
};

template <typename T>
__device__ __forceinline__ void wt_store(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ bool lb_barrier(const ListArgs &a, unsigned target, bool release)
{
    __shared__ int ok_sh;
    const int G = gridDim.x;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (release) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        wt_store(a.sync + LB_ARR0 + blockIdx.x, target);
    }
    if (blockIdx.x == 0 && threadIdx.x < SX_WAVE) {
        bool done = false;
        for (int it = 0; it < LB_SPIN && !done; ++it) {
            bool ok = true;
            for (int b = threadIdx.x; b < G; b += SX_WAVE)
                ok = ok && __hip_atomic_load(a.sync + LB_ARR0 + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target;
            done = __all(ok);
            if (!done) __builtin_amdgcn_s_sleep(1);
        }
        if (!done && threadIdx.x == 0) atomicCAS(a.err, DERR_NONE, DERR_LIST_SYNC);
        if (threadIdx.x < LB_NREL) wt_store(a.sync + LB_REL0 + 32 * threadIdx.x, done ? target : LB_ABORT);
    }
    if (threadIdx.x == 0) {
        const unsigned *r = a.sync + LB_REL0 + 32 * (blockIdx.x % LB_NREL);
        unsigned v = 0;
        int it = 0;
        for (; it < LB_SPIN; ++it) {
            v = __hip_atomic_load(r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (v >= target) break;
            __builtin_amdgcn_s_sleep(1);
        }
        ok_sh = it < LB_SPIN && v != LB_ABORT;
        if (it >= LB_SPIN) atomicCAS(a.err, DERR_NONE, DERR_LIST_SYNC);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    return ok_sh != 0;
}

__device__ __forceinline__ void lb_range(int n, int b, int G, int &i0, int &i1)
{
    i0 = (int)((long)n * b / G);
    i1 = (int)((long)n * (b + 1) / G);
}

// the sorted row by ranks (as rank_row_entries), stored through to memory
template <int PER>
__device__ __forceinline__ void lb_rank_row(const int *sh, int len, int len4, int lane, int q0, int *row)
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
        const int4 v = sh4[m];
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

// The reference's pair test for the build kernel's cell grids, where the
// list radius is below half of every box edge.  Bit for bit the same squared
// distance as in_list_pos without its canonical orientation: IEEE subtraction
// and multiplication are antisymmetric and round-to-nearest-even is
// symmetric, so (pi - pj) and (pj - pi) give the same squares.  The image
// shift rounds to nearest even, which differs from the reference's
// half-away rounding only at an exact half-box separation, where the pair is
// beyond the list radius with either shift.
__device__ __forceinline__ bool lb_pair_in(double4 pi, double4 pj, const DState &s, const CellGrid &g, double rsq_cut)
{
    double d0 = pi.x - pj.x, d1 = pi.y - pj.y, d2 = pi.z - pj.z;
    d0 = d0 - rint(d0 * g.inv[0]) * s.box[0];
    d1 = d1 - rint(d1 * g.inv[1]) * s.box[1];
    d2 = d2 - rint(d2 * g.inv[2]) * s.box[2];
    return d0 * d0 + d1 * d1 + d2 * d2 < rsq_cut;
}

// The candidates that pass the reference's pair test with the cell's
// particle u (returned in i), compacted in candidate order into the LDS
// buffer `row` (the first LB_RB of them); returns their number.  STAGED: the
// candidates are in LDS (cj, cp), every one is tested first and the hits
// compacted after (no LDS stores between the loads, so they pipeline);
// otherwise they are read from the bins through the 27 ranges.
template <bool STAGED>
__device__ __forceinline__ int lb_row_hits(const DState &s, const ListArgs &a, const int *cj, const double4 *cp,
                                           const int *rs, const int *ro, int T, int u, int &i, int *row)
{
    const int lane = threadIdx.x % SX_WAVE;
    // particle u of the cell itself: neighbour range 13 (dx = dy = dz = 0)
    double4 pi;
    if (STAGED) {
        i = cj[ro[13] + u];
        pi = cp[ro[13] + u];
    } else {
        i = a.bin_items[rs[13] + u];
        pi = a.bin_pos[rs[13] + u];
    }
    const unsigned long long below = (1ull << lane) - 1ull;
    int cnt = 0;
    if (STAGED) {
        unsigned long long m[LB_TMAX / SX_WAVE];
#pragma unroll
        for (int k = 0; k < LB_TMAX / SX_WAVE; ++k) {
            m[k] = 0ull;
            if (k * SX_WAVE < T) {
                const int q = k * SX_WAVE + lane;
                bool hit = false;
                if (q < T) hit = cj[q] != i && lb_pair_in(pi, cp[q], s, a.g, a.rsq_cut);
                m[k] = __ballot(hit);
            }
        }
#pragma unroll
        for (int k = 0; k < LB_TMAX / SX_WAVE; ++k) {
            if (k * SX_WAVE >= T) break;
            const int pos = cnt + __popcll(m[k] & below);
            if (((m[k] >> lane) & 1ull) && pos < LB_RB) row[pos] = cj[k * SX_WAVE + lane];
            cnt += __popcll(m[k]);
        }
        return cnt;
    }
    for (int q0 = 0; q0 < T; q0 += SX_WAVE) {
        const int q = q0 + lane;
        int j = -1;
        bool hit = false;
        if (q < T) {
            int r = 0;
            for (int st = 16; st > 0; st >>= 1)
                if (r + st < 27 && ro[r + st] <= q) r += st;
            const long src = rs[r] + q - ro[r];
            j = a.bin_items[src];
            hit = j != i && lb_pair_in(pi, a.bin_pos[src], s, a.g, a.rsq_cut);
        }
        const unsigned long long m = __ballot(hit);
        const int pos = cnt + __popcll(m & below);
        if (hit && pos < LB_RB) row[pos] = j;
        cnt += __popcll(m);
    }
    return cnt;
}

// A row of at most 64 (distinct) entries in LDS sorted into dst: each lane
// ranks its entry against the whole row, every load issued before the
// comparisons
__device__ __forceinline__ void lb_sort_short(const int *sh, int len, int lane, int *dst)
{
    const int mine = lane < len ? sh[lane] : 0x7fffffff;
    const int4 *sh4 = reinterpret_cast<const int4 *>(sh);
    const int n4 = (len + 3) >> 2;
    int4 v[16];
#pragma unroll
    for (int m = 0; m < 16; ++m)
        if (m < n4) v[m] = sh4[m];
    int rank = 0;
#pragma unroll
    for (int m = 0; m < 16; ++m)
        if (m < n4) rank += (v[m].x < mine) + (v[m].y < mine) + (v[m].z < mine) + (v[m].w < mine);
    if (lane < len) dst[rank] = mine;
}

// Stage 5 of k_list_build for the entries of the rows i0 .. i0 + nr - 1,
// whose packed records are ms[0..nr] (in LDS, or in memory); ms[nr] is the
// next row's, whose start ends the range
template <typename MS>
__device__ __forceinline__ void lb_entries(const ListArgs &a, MS ms, int i0, int nr)
{
    const int t = threadIdx.x;
    const int P0 = ms[0].x, P1 = ms[nr].x;
    constexpr int ILP = 8;
    for (int b0 = P0 + t; b0 < P1; b0 += ILP * LB_THREADS) {
        int ie[ILP], je[ILP], tp[ILP], r_[ILP];
#pragma unroll
        for (int u = 0; u < ILP; ++u) {
            const int p = b0 + u * LB_THREADS;
            ie[u] = je[u] = tp[u] = r_[u] = -1;
            if (p >= P1) continue;
            int r = 0, hi = nr;   // the last row start <= p
            while (hi - r > 1) {
                const int mid = (r + hi) >> 1;
                if (ms[mid].x <= p) r = mid; else hi = mid;
            }
            r_[u] = r;
            ie[u] = i0 + r;
            tp[u] = p - ms[r].x;
            je[u] = a.rows[(long)ie[u] * a.rc + tp[u]];
        }
#pragma unroll
        for (int u = 0; u < ILP; ++u)
            if (ie[u] >= 0) a.nbr[b0 + u * LB_THREADS] = je[u];
        if (!a.need_half) continue;
        // upper entries (j > i): the half pair from row i's own record
#pragma unroll
        for (int u = 0; u < ILP; ++u) {
            if (ie[u] < 0 || je[u] < ie[u]) continue;
            const int4 mi = ms[r_[u]];
            const int k = mi.y + (tp[u] - mi.z);
            a.half_i[k] = ie[u];
            a.half_j[k] = je[u];
        }
    }
}

__global__ void __launch_bounds__(LB_THREADS) k_list_build(DState s, ListArgs a, const int *flag)
{
    if (*flag == 0) return;
    __shared__ int rng_s[LB_NQ][27], rng_o[LB_NQ][28];
    __shared__ int cand_j[LB_NQ][LB_TMAX];
    __shared__ double4 cand_p[LB_NQ][LB_TMAX];
    __shared__ __align__(16) int rowbuf[LB_WAVES][LB_RB + 4];
    __shared__ int maxrow_sh, over_sh, batch_sh;
    __shared__ long red[4][LB_WAVES];
    const int n = s.n, G = gridDim.x, t = threadIdx.x, lane = t % SX_WAVE, w = t / SX_WAVE;
    const int tid = blockIdx.x * LB_THREADS + t, nth = G * LB_THREADS;
    const unsigned ep = __hip_atomic_load(a.sync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned base = ep * LB_NBAR;
    // 1. bin, recording the positions the list is built from
    if (tid == 0) {
        if (a.count) a.count[7] += 1;
        wt_store(a.max_row, 0);
    }
    for (int i = tid; i < n; i += nth) {
        const double x = s.f[SX_RX][i], y = s.f[SX_RY][i], z = s.f[SX_RZ][i];
        a.xl[i] = x; a.yl[i] = y; a.zl[i] = z;
        const int cx = cell_coord(x, s.box[0], a.g.nc[0], a.g.w[0]);
        const int cy = cell_coord(y, s.box[1], a.g.nc[1], a.g.w[1]);
        const int cz = cell_coord(z, s.box[2], a.g.nc[2], a.g.w[2]);
        const int c = (cx * a.g.nc[1] + cy) * a.g.nc[2] + cz;
        const int slot = atomicAdd(&a.cell_count[c], 1);
        if (slot < a.cap) {
            const long q = (long)c * a.cap + slot;
            a.bin_items[q] = i;
            a.bin_pos[q] = make_double4(x, y, z, 0.0);
        } else {
            atomicMax(a.max_cell, slot + 1);
        }
    }
    if (!lb_barrier(a, base + 1, true)) return;
    if (blockIdx.x == 0 && t == 0) wt_store(a.sync, ep + 1u);   // every workgroup has read the epoch
    // 2. rows, LB_NQ cells at a time: quarter h of the wavefronts takes cell c0 + h
    if (t == 0) maxrow_sh = 0;
    const int nx = a.g.nc[0], ny = a.g.nc[1], nz = a.g.nc[2];
    const int h = w / LB_QW, ht = t % LB_QT, hw = w % LB_QW;
    int *myrow = rowbuf[w];
    unsigned *ctr = a.sync + LB_CTR0 + 32 * (ep & 1u);
    for (;;) {
        // the next batch of LB_NQ cells (batches handed out in order)
        if (t == 0) batch_sh = (int)atomicAdd(ctr, 1u);
        __syncthreads();
        const int c0 = LB_NQ * batch_sh;
        if (c0 >= a.ncells) break;
        const int c = c0 + h;
        const bool live = c < a.ncells;
        int len = 0;
        if (live && ht < 27) {
            const int cz = c % nz, cy = (c / nz) % ny, cx = c / (nz * ny);
            const int x = (cx + ht / 9 - 1 + nx) % nx, y = (cy + (ht / 3) % 3 - 1 + ny) % ny;
            const int z = (cz + ht % 3 - 1 + nz) % nz;
            const int cc = (x * ny + y) * nz + z;
            rng_s[h][ht] = cc * a.cap;
            len = min(a.cell_count[cc], a.cap);
        }
        if (hw == 0) {   // offsets of the 27 ranges
            int x = len;
            for (int off = 1; off < 32; off <<= 1) {
                const int y = __shfl_up(x, off, SX_WAVE);
                if (lane >= off) x += y;
            }
            if (lane < 27) rng_o[h][lane + 1] = x;
            if (lane == 0) rng_o[h][0] = 0;
        }
        __syncthreads();
        const int T = live ? rng_o[h][27] : 0;
        const bool staged = T <= LB_TMAX;
        if (staged)
            for (int q = ht; q < T; q += LB_QT) {
                int r = 0;
                for (int st = 16; st > 0; st >>= 1)
                    if (r + st < 27 && rng_o[h][r + st] <= q) r += st;
                const long src = rng_s[h][r] + q - rng_o[h][r];
                cand_j[h][q] = a.bin_items[src];
                cand_p[h][q] = a.bin_pos[src];
            }
        __syncthreads();
        const int own = live ? rng_o[h][14] - rng_o[h][13] : 0;
        for (int u = hw; u < own; u += LB_QW) {
            int i;
            const int cnt = staged ? lb_row_hits<true>(s, a, cand_j[h], cand_p[h], rng_s[h], rng_o[h], T, u, i, myrow)
                                   : lb_row_hits<false>(s, a, cand_j[h], cand_p[h], rng_s[h], rng_o[h], T, u, i, myrow);
            int low = 0;
            if (cnt <= LB_RB && cnt <= a.rc) {
                // each entry's rank among the row's (distinct) entries is its
                // place in the sorted row (padded to a multiple of 4)
                const int cnt4 = (cnt + 3) & ~3;
                for (int e = cnt + lane; e < cnt4; e += SX_WAVE) myrow[e] = 0x7fffffff;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                int *dst = a.rows + (long)i * a.rc;
                if (cnt <= SX_WAVE) {
                    lb_sort_short(myrow, cnt, lane, dst);
                } else {
                    const int per = min(SORT_PER_LANE, (cnt + SX_WAVE - 1) / SX_WAVE);
                    for (int q0 = 0; q0 < cnt; q0 += SX_WAVE * per) {
                        switch (per) {
                        case 1: lb_rank_row<1>(myrow, cnt, cnt4, lane, q0, dst); break;
                        case 2: lb_rank_row<2>(myrow, cnt, cnt4, lane, q0, dst); break;
                        case 3: lb_rank_row<3>(myrow, cnt, cnt4, lane, q0, dst); break;
                        case 4: lb_rank_row<4>(myrow, cnt, cnt4, lane, q0, dst); break;
                        case 5: lb_rank_row<5>(myrow, cnt, cnt4, lane, q0, dst); break;
                        case 6: lb_rank_row<6>(myrow, cnt, cnt4, lane, q0, dst); break;
                        case 7: lb_rank_row<7>(myrow, cnt, cnt4, lane, q0, dst); break;
                        default: lb_rank_row<8>(myrow, cnt, cnt4, lane, q0, dst); break;
                        }
                    }
                }
                for (int e = lane; e < cnt; e += SX_WAVE) low += myrow[e] < i;
            }
            for (int off = SX_WAVE / 2; off > 0; off >>= 1) low += __shfl_xor(low, off, SX_WAVE);
            if (lane == 0) {
                a.row_len[i] = cnt;
                a.lower[i] = low;
                a.half_len[i] = cnt - low;
                atomicMax(&maxrow_sh, cnt);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        __syncthreads();   // the staged candidates are replaced next
    }
    if (t == 0) atomicMax(a.max_row, maxrow_sh);
    if (!lb_barrier(a, base + 2, true)) return;
    // 3. sums of the row lengths of this workgroup's particles; the cell
    // counts and this epoch's batch counter back to zero for later builds
    for (int c = tid; c < a.ncells; c += nth) a.cell_count[c] = 0;
    if (tid == 0) wt_store(ctr, 0u);
    int i0, i1;
    lb_range(n, blockIdx.x, G, i0, i1);
    {
        long f = 0, hh = 0;
        for (int i = i0 + t; i < i1; i += LB_THREADS) {
            f += a.row_len[i];
            hh += a.half_len[i];
        }
        for (int off = SX_WAVE / 2; off > 0; off >>= 1) {
            f += __shfl_xor(f, off, SX_WAVE);
            hh += __shfl_xor(hh, off, SX_WAVE);
        }
        if (lane == 0) { red[0][w] = f; red[1][w] = hh; }
        __syncthreads();
        if (t == 0) {
            long F = 0, H = 0;
            for (int q = 0; q < LB_WAVES; ++q) { F += red[0][q]; H += red[1][q]; }
            wt_store(a.bsum + 2 * blockIdx.x, (int)F);
            wt_store(a.bsum + 2 * blockIdx.x + 1, (int)H);
        }
    }
    if (!lb_barrier(a, base + 3, false)) return;
    // 4. offsets, totals and the starts of this workgroup's rows
    {
        long fb = 0, hb = 0, ft = 0, ht2 = 0;
        for (int b = t; b < G; b += LB_THREADS) {
            const long f = a.bsum[2 * b], hh = a.bsum[2 * b + 1];
            if (b < (int)blockIdx.x) { fb += f; hb += hh; }
            ft += f;
            ht2 += hh;
        }
        for (int off = SX_WAVE / 2; off > 0; off >>= 1) {
            fb += __shfl_xor(fb, off, SX_WAVE);
            hb += __shfl_xor(hb, off, SX_WAVE);
            ft += __shfl_xor(ft, off, SX_WAVE);
            ht2 += __shfl_xor(ht2, off, SX_WAVE);
        }
        __syncthreads();
        if (lane == 0) { red[0][w] = fb; red[1][w] = hb; red[2][w] = ft; red[3][w] = ht2; }
        __syncthreads();
        fb = hb = ft = ht2 = 0;
        for (int q = 0; q < LB_WAVES; ++q) { fb += red[0][q]; hb += red[1][q]; ft += red[2][q]; ht2 += red[3][q]; }
        // a cell beyond its room or a row beyond the scratch stride or the
        // LDS row buffer leaves the list empty too
        const int mr = __hip_atomic_load(a.max_row, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int mc = __hip_atomic_load(a.max_cell, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const bool roomerr = mr > min(LB_RB, a.rc) || mc > a.cap;
        const bool over = ft > a.cap_full || roomerr;   // uniform: every workgroup sees the same sums
        if (t == 0) {
            over_sh = over;
            if (blockIdx.x == 0) {
                *a.total = ft > 0x7fffffffL ? 0x7fffffff : (int)ft;
                *a.total_half = over ? 0 : (int)ht2;
                if (over) {
                    atomicCAS(a.err, DERR_NONE, roomerr ? DERR_ROW_CAP : DERR_NLIST_CAP);
                } else {
                    a.row_start[n] = (int)ft;
                    if (a.need_half) a.half_start[n] = (int)ht2;
                    a.meta[n] = make_int4((int)ft, (int)ht2, 0, 0);
                }
            }
        }
        if (over) {   // an empty list (as the stage kernels leave one)
            for (int i = i0 + t; i < i1; i += LB_THREADS) {
                a.row_start[i] = 0;
                a.half_len[i] = 0;
                if (a.need_half) a.half_start[i] = 0;
            }
            if (blockIdx.x == 0 && t == 0) {
                a.row_start[n] = 0;
                if (a.need_half) a.half_start[n] = 0;
            }
        } else if (w == 0) {
            long cf = fb, ch = hb;
            for (int b0 = i0; b0 < i1; b0 += SX_WAVE) {
                const int i = b0 + lane;
                const int f = i < i1 ? a.row_len[i] : 0, hh = i < i1 ? a.half_len[i] : 0;
                int xf = f, xh = hh;
                for (int off = 1; off < SX_WAVE; off <<= 1) {
                    const int yf = __shfl_up(xf, off, SX_WAVE), yh = __shfl_up(xh, off, SX_WAVE);
                    if (lane >= off) { xf += yf; xh += yh; }
                }
                if (i < i1) {
                    const int rsi = (int)(cf + xf - f), hsi = (int)(ch + xh - hh);
                    a.row_start[i] = rsi;
                    if (a.need_half) a.half_start[i] = hsi;
                    a.meta[i] = make_int4(rsi, hsi, f - hh, f);
                }
                cf += __shfl(xf, SX_WAVE - 1, SX_WAVE);
                ch += __shfl(xh, SX_WAVE - 1, SX_WAVE);
            }
        }
    }
    __syncthreads();
    const bool over = over_sh != 0;
    if (!lb_barrier(a, base + 4, true)) return;
    if (over) return;
    // 5. rows into the list and the half-pair index, over this workgroup's
    // entries [row_start[i0], row_start[i1]); the row of an entry from a
    // search of the range's row starts (in LDS when they fit)
    const int nr = i1 - i0;
    int4 *ms_sh = reinterpret_cast<int4 *>(&cand_p[0][0]);
    if (nr + 1 <= LB_NQ * LB_TMAX * 2) {
        for (int q = t; q <= nr; q += LB_THREADS) ms_sh[q] = a.meta[i0 + q];
        __syncthreads();
        lb_entries(a, ms_sh, i0, nr);
    } else {
        lb_entries(a, (const int4 *)(a.meta + i0), i0, nr);
    }
}

static int lb_blocks_per_cu()
{
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_list_build, LB_THREADS, 0) != hipSuccess) return 0;
    return per;
}

// workgroups of the one-launch rebuild: one per CU (every one of them
// resident at once), 0 when a workgroup does not fit a CU
static int lb_grid_for()
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return lb_blocks_per_cu() >= 1 ? std::min(cus, LB_SYNC_WORDS - LB_ARR0) : 0;
}

static void launch_list_rebuild(sx_ctx *c, const int *flag)
{
    NList &nl = c->nl;
    ListArgs a;
    a.g = cell_grid(c);
    a.cell_count = nl.cell_count;
    a.bin_items = nl.lb_bin_items;
    a.bin_pos = (double4 *)nl.lb_bin_pos;
    a.cap = nl.lb_cap;
    a.xl = nl.x_last; a.yl = nl.y_last; a.zl = nl.z_last;
    a.count = c->d_counts;
    a.ncells = nl.ncells;
    a.rsq_cut = nl.r_skin_sq;
    a.row_len = nl.row_len; a.lower = nl.lower; a.half_len = nl.half_len; a.row_start = nl.row_start; a.nbr = nl.nbr;
    a.cap_full = nl.cap_full;
    a.need_half = nl.need_half ? 1 : 0;
    a.half_start = nl.half_start; a.half_i = nl.half_i; a.half_j = nl.half_j;
    a.total = c->d_flag + 1; a.total_half = c->d_flag + 2; a.max_row = c->d_flag + 3; a.max_cell = c->d_flag + 5;
    a.rows = nl.lb_rows;
    a.rc = nl.lb_rc;
    a.meta = (int4 *)nl.lb_meta;
    a.bsum = nl.lb_bsum;
    a.err = c->d_err;
    a.sync = nl.lb_sync;
    hipLaunchKernelGGL(k_list_build, dim3(nl.lb_grid), dim3(LB_THREADS), 0, c->stream, c->dstate(), a, flag);
}

// after DERR_LIST_SYNC: a build whose workgroups gave up waiting left its
// arrival count mid-build
void sx_nlist_reset_sync(sx_ctx *c)
{
    NList &nl = c->nl;
    if (nl.lb_sync) HIPCHK(hipMemsetAsync(nl.lb_sync, 0, sizeof(unsigned) * LB_SYNC_WORDS, c->stream));
    if (nl.ncells > 0) HIPCHK(hipMemsetAsync(nl.cell_count, 0, sizeof(int) * (nl.ncells + 1), c->stream));
}

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
                       nl.nbr, nl.half_i, nl.half_j, flag, nl.cap_half, c->d_err)
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
                 nl.half_start, nl.half_len, nl.nbr, nl.half_i, nl.half_j};
    for (int *p : ip)
        if (p) hipFree(p);
    double *dp[] = {nl.x_last, nl.y_last, nl.z_last, nl.cell_pos};
    for (double *p : dp)
        if (p) hipFree(p);
    nl.cell_of = nl.cell_count = nl.cell_start = nl.cell_items = nl.row_len = nl.row_start = nl.lower = nullptr;
    nl.half_start = nl.half_len = nl.nbr = nl.half_i = nl.half_j = nullptr;
    nl.x_last = nl.y_last = nl.z_last = nl.cell_pos = nullptr;
    if (nl.lb_sync) hipFree(nl.lb_sync);
    if (nl.lb_rows) hipFree(nl.lb_rows);
    if (nl.lb_bsum) hipFree(nl.lb_bsum);
    if (nl.lb_bin_items) hipFree(nl.lb_bin_items);
    if (nl.lb_bin_pos) hipFree(nl.lb_bin_pos);
    if (nl.lb_meta) hipFree(nl.lb_meta);
    nl.lb_meta = nullptr;
    nl.lb_sync = nullptr;
    nl.lb_rows = nl.lb_bsum = nl.lb_bin_items = nullptr;
    nl.lb_bin_pos = nullptr;
    nl.lb_rc = nl.lb_cap = 0;
    nl.lb_grid = 0;
}

static void nl_alloc_lists(sx_ctx *c, long cap_full)
{
    NList &nl = c->nl;
    if (nl.nbr) hipFree(nl.nbr);
    if (nl.half_i) hipFree(nl.half_i);
    if (nl.half_j) hipFree(nl.half_j);
    if (cap_full > 0x7fffffffL) throw sx_error{SX_ERR_CAPACITY, "neighbour list larger than 2^31 entries"};
    nl.cap_full = cap_full;
    nl.cap_half = cap_full / 2 + 1;
    HIPCHK(hipMalloc(&nl.nbr, sizeof(int) * nl.cap_full));
    HIPCHK(hipMalloc(&nl.half_i, sizeof(int) * nl.cap_half));
    HIPCHK(hipMalloc(&nl.half_j, sizeof(int) * nl.cap_half));
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
    int **full[] = {&nl.nbr};
    int **half[] = {&nl.half_i, &nl.half_j};
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
    // rows of the one-launch rebuild: room for half as long again
    if (nl.lb_grid > 0) {
        const int mr = nl_max_row(c);
        if (3 * mr / 2 + 32 > nl.lb_rc && !lb_scratch(c, 2 * mr + 64)) nl.lb_grid = 0;
    }
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
    if (e == DERR_ROW_CAP && nl.init && nl.lb_grid > 0) {
        // a row (d_flag[3]: the longest) or a cell (d_flag[5]: the fullest,
        // when one overflowed) of a replayed rebuild outgrew its room; the
        // list of that rebuild was left empty
        int f[6] = {0, 0, 0, 0, 0, 0};
        HIPCHK(hipMemcpyAsync(f, c->d_flag, sizeof(int) * 6, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        const int need_row = f[3], need_cell = f[5], had_row = nl.lb_rc, had_cell = nl.lb_cap;
        HIPCHK(hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream));
        HIPCHK(hipMemsetAsync(c->d_flag + 5, 0, sizeof(int), c->stream));
        if (need_row > had_row && !lb_scratch(c, 2 * need_row + 64)) nl.lb_grid = 0;
        if (need_cell > had_cell) lb_bins(c, 2 * need_cell + 32);
        sx_nlist_build_impl(c, false);
        throw sx_error{SX_ERR_CAPACITY,
                       "neighbour row or cell larger than the rebuild's room during the run (row " +
                           std::to_string(need_row) + " entries, room for " + std::to_string(had_row) + "; cell " +
                           std::to_string(need_cell) + " particles, room for " + std::to_string(had_cell) +
                           "): from that rebuild on the list was empty and the steps ran without pair forces; the "
                           "room has been grown and the list rebuilt at the current positions"};
    }
    sx_check_device_error(c);
    if (headroom) sx_nlist_headroom(c);
}

// scratch rows of the one-launch rebuild with stride rc (at most the LDS row
// buffer); false when rows that long do not fit it (the stage kernels build
// then)
static bool lb_scratch(sx_ctx *c, int rc)
{
    NList &nl = c->nl;
    rc = (std::max(rc, 64) + 3) & ~3;   // rows 16-byte aligned
    if (rc > LB_RB) return false;
    if (nl.lb_rows && rc <= nl.lb_rc) return true;
    if (nl.lb_rows) hipFree(nl.lb_rows);
    nl.lb_rows = nullptr;
    HIPCHK(hipMalloc(&nl.lb_rows, sizeof(int) * (size_t)c->n * rc));
    nl.lb_rc = rc;
    c->have_graph = false;
    return true;
}

// cell bins of the one-launch rebuild with room for cap particles per cell
static void lb_bins(sx_ctx *c, int cap)
{
    NList &nl = c->nl;
    cap = std::max(cap, 16);
    if (nl.lb_bin_items && cap <= nl.lb_cap) return;
    if (nl.lb_bin_items) hipFree(nl.lb_bin_items);
    if (nl.lb_bin_pos) hipFree(nl.lb_bin_pos);
    nl.lb_bin_items = nullptr;
    nl.lb_bin_pos = nullptr;
    HIPCHK(hipMalloc(&nl.lb_bin_items, sizeof(int) * (size_t)nl.ncells * cap));
    HIPCHK(hipMalloc(&nl.lb_bin_pos, sizeof(double) * 4 * (size_t)nl.ncells * cap));
    nl.lb_cap = cap;
    c->have_graph = false;
}

__global__ void k_max_row(const int *row_len, int n, int *out)
{
    int m = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = max(m, row_len[i]);
    for (int off = SX_WAVE / 2; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, SX_WAVE));
    if ((threadIdx.x & (SX_WAVE - 1)) == 0) atomicMax(out, m);
}

// the longest row of the current list (d_flag[3])
static int nl_max_row(sx_ctx *c)
{
    int m = 0;
    HIPCHK(hipMemsetAsync(c->d_flag + 3, 0, sizeof(int), c->stream));
    hipLaunchKernelGGL(k_max_row, dim3(std::min(sx_blocks(c->n, 256), 256)), dim3(256), 0, c->stream, c->nl.row_len,
                       c->n, c->d_flag + 3);
    HIPCHK(hipMemcpyAsync(&m, c->d_flag + 3, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return m;
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
    // the one-launch conditional rebuild (cell grids only), its scratch rows
    // sized for twice the expected row
    if (nl.use_cells) {
        nl.lb_grid = lb_grid_for();
        if (nl.lb_grid > 0 && !lb_scratch(c, (int)std::ceil(2.0 * row_est) + 64)) nl.lb_grid = 0;
        if (nl.lb_grid > 0) {
            HIPCHK(hipMalloc(&nl.lb_sync, sizeof(unsigned) * LB_SYNC_WORDS));
            HIPCHK(hipMemsetAsync(nl.lb_sync, 0, sizeof(unsigned) * LB_SYNC_WORDS, c->stream));
            HIPCHK(hipMalloc(&nl.lb_bsum, sizeof(int) * 2 * nl.lb_grid));
            HIPCHK(hipMalloc(&nl.lb_meta, sizeof(int) * 4 * ((size_t)n + 1)));
            lb_bins(c, (int)std::ceil(2.0 * n / nl.ncells) + 32);
        }
    }
    nl.init = true;
}

// the reference's half list (half_start, half_i, half_j; d_flag[2] = its
// length) and the pair positions, from the sorted rows
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
    if (conditional && nl.use_cells && nl.lb_grid > 0 && cell_grid(c).fast) {
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
