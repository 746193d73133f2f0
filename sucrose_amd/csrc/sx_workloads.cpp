// Seeded synthetic initial states for the benchmark workloads (host code).
//
//  sx_gen_ellipsoids_rsa: random sequential addition of identical ellipsoids
//    with uniformly random orientations in a cubic periodic box; a candidate
//    is accepted when it overlaps no placed ellipsoid, overlap being decided
//    by the Perram-Wertheim contact function F = max_l l(1-l) r^T G(l)^-1 r
//    (overlap iff F < 1), the function whose maximiser the ECP kernels solve
//    for (force.f90:1188-1200).
//  sx_gen_uniform_quaternions: uniformly distributed unit quaternions
//    (Shoemake's subgroup algorithm from three uniforms).
//  sx_gen_chains: self-avoiding random-walk chains of fixed bond length in a
//    cubic periodic box (a bead is rejected closer than min_dist to any
//    placed bead; a chain that cannot grow is withdrawn and restarted).
//  Random numbers come from the reference's Philox4x32-10 stream, counter
//  (seed, attempt, k, 0x57524b4c).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/sucrose.h"

namespace {

struct Rng {
    uint32_t seed;
    uint64_t ctr = 0;
    uint32_t buf[4];
    int left = 0;
    double u()
    {
        if (left == 0) {
            uint32_t c[4] = {seed, (uint32_t)ctr, (uint32_t)(ctr >> 32), 0x57524b4cu};
            sx_philox4x32_10(c, buf);
            ++ctr;
            left = 4;
        }
        return (double)buf[--left] * (1.0 / 4294967296.0);
    }
};

void chol_solve(const double g[6], const double b[3], double o[3])
{
    // g: a11 a21 a31 a22 a32 a33
    double i11 = 1.0 / std::sqrt(g[0]);
    double l21 = g[1] * i11, l31 = g[2] * i11;
    double i22 = 1.0 / std::sqrt(g[3] - l21 * l21);
    double l32 = (g[4] - l31 * l21) * i22;
    double i33 = 1.0 / std::sqrt(g[5] - l31 * l31 - l32 * l32);
    double y1 = b[0] * i11, y2 = (b[1] - l21 * y1) * i22, y3 = (b[2] - l31 * y1 - l32 * y2) * i33;
    o[2] = y3 * i33;
    o[1] = (y2 - l32 * o[2]) * i22;
    o[0] = (y1 - l21 * o[1] - l31 * o[2]) * i11;
}

double pwF(double l, const double *A, const double *B, const double r[3])
{
    double g[6], k[3];
    for (int q = 0; q < 6; ++q) g[q] = (1 - l) * A[q] + l * B[q];
    chol_solve(g, r, k);
    return l * (1 - l) * (r[0] * k[0] + r[1] * k[1] + r[2] * k[2]);
}

// true if the ellipsoids overlap (max_l F < 1)
bool overlap(const double *A, const double *B, const double r[3])
{
    const double gr = 0.6180339887498949;
    double a = 0.0, b = 1.0, c = b - gr * (b - a), d = a + gr * (b - a);
    double fc = pwF(c, A, B, r), fd = pwF(d, A, B, r);
    if (fc >= 1.0 || fd >= 1.0) return false;
    for (int it = 0; it < 60 && (b - a) > 1e-12; ++it) {
        if (fc > fd) { b = d; d = c; fd = fc; c = b - gr * (b - a); fc = pwF(c, A, B, r); }
        else { a = c; c = d; fc = fd; d = a + gr * (b - a); fd = pwF(d, A, B, r); }
        if (fc >= 1.0 || fd >= 1.0) return false;
    }
    return true;
}

// shape matrix R diag(s^2) R^T (lower triangle) for quaternion (w,x,y,z)
void shape(const double q[4], const double ax[3], double A[6])
{
    double w = q[0], x = q[1], y = q[2], z = q[3];
    double R[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)},
                      {2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)},
                      {2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)}};
    const int ij[6][2] = {{0, 0}, {1, 0}, {2, 0}, {1, 1}, {2, 1}, {2, 2}};
    for (int t = 0; t < 6; ++t) {
        int i = ij[t][0], j = ij[t][1];
        double s = 0;
        for (int k = 0; k < 3; ++k) s += R[i][k] * ax[k] * ax[k] * R[j][k];
        A[t] = s;
    }
}

}  // namespace

extern "C" {

// Places up to n ellipsoids with semi-axes (a, b, c) in a periodic cube of
// edge L.  pos[3n], quat[4n] (w x y z).  Returns the number placed.
int sx_gen_ellipsoids_rsa(int n, double a, double b, double c, double L, uint32_t seed, long max_attempts,
                          double *pos, double *quat)
{
    Rng rng{seed};
    double amax = std::max(a, std::max(b, c)), amin = std::min(a, std::min(b, c));
    double ax[3] = {a, b, c};
    int nc = std::max(1, (int)std::floor(L / (2 * amax)));
    if (nc < 3) nc = 1;
    double cw = L / nc;
    std::vector<int> head((size_t)nc * nc * nc, -1), next(n, -1);
    std::vector<double> shp(6 * (size_t)n);
    int placed = 0;
    for (long att = 0; att < max_attempts && placed < n; ++att) {
        double p[3] = {(rng.u() - 0.5) * L, (rng.u() - 0.5) * L, (rng.u() - 0.5) * L};
        double u1 = rng.u(), u2 = rng.u(), u3 = rng.u();
        const double tp = 6.283185307179586;
        double q[4] = {std::sqrt(1 - u1) * std::sin(tp * u2), std::sqrt(1 - u1) * std::cos(tp * u2),
                       std::sqrt(u1) * std::sin(tp * u3), std::sqrt(u1) * std::cos(tp * u3)};
        double A[6];
        shape(q, ax, A);
        int ci[3];
        for (int k = 0; k < 3; ++k) {
            ci[k] = (int)std::floor((p[k] + L / 2) / cw);
            ci[k] = std::min(std::max(ci[k], 0), nc - 1);
        }
        bool bad = false;
        int span = nc >= 3 ? 1 : 0;
        for (int dx = -span; dx <= span && !bad; ++dx)
            for (int dy = -span; dy <= span && !bad; ++dy)
                for (int dz = -span; dz <= span && !bad; ++dz) {
                    int cx = (ci[0] + dx + nc) % nc, cy = (ci[1] + dy + nc) % nc, cz = (ci[2] + dz + nc) % nc;
                    for (int j = head[((size_t)cx * nc + cy) * nc + cz]; j >= 0 && !bad; j = next[j]) {
                        double r[3];
                        for (int k = 0; k < 3; ++k) {
                            r[k] = p[k] - pos[3 * j + k];
                            r[k] -= std::round(r[k] / L) * L;
                        }
                        double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                        if (r2 >= 4 * amax * amax) continue;
                        if (r2 < 4 * amin * amin || overlap(A, &shp[6 * (size_t)j], r)) bad = true;
                    }
                }
        if (bad) continue;
        for (int k = 0; k < 3; ++k) pos[3 * placed + k] = p[k];
        for (int k = 0; k < 4; ++k) quat[4 * placed + k] = q[k];
        for (int k = 0; k < 6; ++k) shp[6 * (size_t)placed + k] = A[k];
        size_t cell = ((size_t)ci[0] * nc + ci[1]) * nc + ci[2];
        next[placed] = head[cell];
        head[cell] = placed;
        ++placed;
    }
    return placed;
}

// Count of overlapping pairs of identical ellipsoids (all-pairs within 2 amax, cell list).
long sx_count_ellipsoid_overlaps(int n, double a, double b, double c, double L, const double *pos, const double *quat)
{
    double amax = std::max(a, std::max(b, c));
    double ax[3] = {a, b, c};
    std::vector<double> shp(6 * (size_t)n);
    for (int i = 0; i < n; ++i) shape(&quat[4 * i], ax, &shp[6 * (size_t)i]);
    int nc = std::max(1, (int)std::floor(L / (2 * amax)));
    if (nc < 3) nc = 1;
    double cw = L / nc;
    std::vector<int> head((size_t)nc * nc * nc, -1), next(n, -1);
    std::vector<int> cell(n);
    for (int i = 0; i < n; ++i) {
        int ci[3];
        for (int k = 0; k < 3; ++k) {
            double x = pos[3 * i + k] - std::floor(pos[3 * i + k] / L + 0.5) * L;
            ci[k] = std::min(std::max((int)std::floor((x + L / 2) / cw), 0), nc - 1);
        }
        cell[i] = (ci[0] * nc + ci[1]) * nc + ci[2];
        next[i] = head[cell[i]];
        head[cell[i]] = i;
    }
    long cnt = 0;
    int span = nc >= 3 ? 1 : 0;
    for (int i = 0; i < n; ++i) {
        int c0 = cell[i];
        int cz = c0 % nc, cy = (c0 / nc) % nc, cx = c0 / (nc * nc);
        for (int dx = -span; dx <= span; ++dx)
            for (int dy = -span; dy <= span; ++dy)
                for (int dz = -span; dz <= span; ++dz) {
                    int x = (cx + dx + nc) % nc, y = (cy + dy + nc) % nc, z = (cz + dz + nc) % nc;
                    for (int j = head[((size_t)x * nc + y) * nc + z]; j >= 0; j = next[j]) {
                        if (j <= i) continue;
                        double r[3];
                        for (int k = 0; k < 3; ++k) {
                            r[k] = pos[3 * i + k] - pos[3 * j + k];
                            r[k] -= std::round(r[k] / L) * L;
                        }
                        double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                        if (r2 >= 4 * amax * amax) continue;
                        if (overlap(&shp[6 * (size_t)i], &shp[6 * (size_t)j], r)) ++cnt;
                    }
                }
    }
    return cnt;
}

// n uniformly distributed unit quaternions (w x y z)
void sx_gen_uniform_quaternions(int n, uint32_t seed, double *quat)
{
    Rng rng{seed};
    const double tp = 6.283185307179586;
    for (int i = 0; i < n; ++i) {
        double u1 = rng.u(), u2 = rng.u(), u3 = rng.u();
        quat[4 * i + 0] = std::sqrt(1 - u1) * std::sin(tp * u2);
        quat[4 * i + 1] = std::sqrt(1 - u1) * std::cos(tp * u2);
        quat[4 * i + 2] = std::sqrt(u1) * std::sin(tp * u3);
        quat[4 * i + 3] = std::sqrt(u1) * std::cos(tp * u3);
    }
}

// nchains self-avoiding random walks of nbeads beads, bond length `bond`, no
// two beads closer than min_dist (< bond), in a periodic cube of edge L;
// pos[3 * nchains * nbeads] wrapped into [-L/2, L/2).  Returns the number of
// complete chains placed.
int sx_gen_chains(int nchains, int nbeads, double L, double bond, double min_dist, uint32_t seed,
                  long max_attempts, double *pos)
{
    Rng rng{seed};
    const long n = (long)nchains * nbeads;
    int nc = std::max(1, (int)std::floor(L / std::max(min_dist, 1e-9)));
    if (nc < 3) nc = 1;
    double cw = L / nc;
    std::vector<long> head((size_t)nc * nc * nc, -1), next((size_t)n, -1), prev((size_t)n, -1);
    std::vector<size_t> cellof((size_t)n);
    auto wrap = [&](double x) { return x - std::floor(x / L + 0.5) * L; };
    auto cell_of = [&](const double *p) {
        int ci[3];
        for (int k = 0; k < 3; ++k) ci[k] = std::min(std::max((int)std::floor((p[k] + L / 2) / cw), 0), nc - 1);
        return ((size_t)ci[0] * nc + ci[1]) * nc + ci[2];
    };
    auto clash = [&](const double *p, long skip) {
        int c0 = (int)cell_of(p);
        int cz = c0 % nc, cy = (c0 / nc) % nc, cx = c0 / (nc * nc);
        int span = nc >= 3 ? 1 : 0;
        for (int dx = -span; dx <= span; ++dx)
            for (int dy = -span; dy <= span; ++dy)
                for (int dz = -span; dz <= span; ++dz) {
                    size_t c = ((size_t)((cx + dx + nc) % nc) * nc + (cy + dy + nc) % nc) * nc + (cz + dz + nc) % nc;
                    for (long j = head[c]; j >= 0; j = next[j]) {
                        if (j == skip) continue;
                        double r2 = 0;
                        for (int k = 0; k < 3; ++k) {
                            double d = p[k] - pos[3 * j + k];
                            d -= std::round(d / L) * L;
                            r2 += d * d;
                        }
                        if (r2 < min_dist * min_dist) return true;
                    }
                }
        return false;
    };
    auto insert = [&](long i) {
        size_t c = cell_of(&pos[3 * i]);
        cellof[i] = c;
        next[i] = head[c];
        prev[i] = -1;
        if (head[c] >= 0) prev[head[c]] = i;
        head[c] = i;
    };
    auto remove = [&](long i) {
        size_t c = cellof[i];
        if (prev[i] >= 0) next[prev[i]] = next[i]; else head[c] = next[i];
        if (next[i] >= 0) prev[next[i]] = prev[i];
    };
    long attempts = 0;
    int done = 0;
    for (int ch = 0; ch < nchains && attempts < max_attempts; ++ch) {
        long base = (long)ch * nbeads;
        int b = 0;
        int stuck = 0;
        while (b < nbeads && attempts < max_attempts) {
            ++attempts;
            double p[3];
            if (b == 0) {
                for (int k = 0; k < 3; ++k) p[k] = (rng.u() - 0.5) * L;
            } else {
                double z = 2 * rng.u() - 1, ph = 6.283185307179586 * rng.u(), s = std::sqrt(1 - z * z);
                const double *q = &pos[3 * (base + b - 1)];
                p[0] = wrap(q[0] + bond * s * std::cos(ph));
                p[1] = wrap(q[1] + bond * s * std::sin(ph));
                p[2] = wrap(q[2] + bond * z);
            }
            if (clash(p, b > 0 ? base + b - 1 : -1)) {
                if (++stuck > 200) {   // withdraw the chain and start it again
                    for (int k = 0; k < b; ++k) remove(base + k);
                    b = 0;
                    stuck = 0;
                }
                continue;
            }
            for (int k = 0; k < 3; ++k) pos[3 * (base + b) + k] = p[k];
            insert(base + b);
            ++b;
            stuck = 0;
        }
        if (b == nbeads) ++done; else break;
    }
    return done;
}

}  // extern "C"
