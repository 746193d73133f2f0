/*
 * CPU oracle for the SUCROSE force-and-integrate path -- TEST INFRASTRUCTURE.
 * An FP64 restatement of the reference routines, written to keep each
 * expression's evaluation order so that it agrees bit-for-bit with the
 * reference built from source (compile with -ffp-contract=off).
 * Only tests, the smoke check and the benchmark's CPU-baseline leg use it.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_REF 3.1415926535897932384626433832795 /* xmath.f90:120 (-r8) */

static inline double anint(double x) { return round(x); }

/* RIJ - ANINT(RIJ/box)*box, componentwise (e.g. system.f90:945) */
static inline double mimg(double d, double L) { return d - anint(d / L) * L; }

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (xmath.f90:14-76)                                     */
/* ------------------------------------------------------------------ */
void or_philox4x32_10_key(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u; /* xmath.f90:5-6 */
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u; /* xmath.f90:7-8 */
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 1; round <= 10; ++round) {
        uint64_t p0 = (uint64_t)M0 * c0;
        uint64_t p1 = (uint64_t)M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ (c1 ^ k0);
        c1 = lo1;
        c2 = hi0 ^ (c3 ^ k1);
        c3 = lo0;
        if (round < 10) { k0 += W0; k1 += W1; } /* wrapping int32 add */
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void or_philox4x32_10(const uint32_t ctr[4], uint32_t out[4])
{
    static const uint32_t key[2] = {0x6D2B79F5u, 0x1B873593u}; /* xmath.f90:10-11 */
    or_philox4x32_10_key(ctr, key, out);
}

/* philox_real: word * 2^-32 (xmath.f90:28-36) */
static inline double philox_real(uint32_t w) { return (double)w * (1.0 / 4294967296.0); }

void or_philox_uniform(const int32_t ctr[4], double out[4])
{
    uint32_t c[4] = {(uint32_t)ctr[0], (uint32_t)ctr[1], (uint32_t)ctr[2], (uint32_t)ctr[3]};
    uint32_t w[4];
    or_philox4x32_10(c, w);
    for (int k = 0; k < 4; ++k) out[k] = philox_real(w[k]);
}

/* philox_normal_3seq (xmath.f90:78-113) */
void or_philox_normal_3seq(int32_t c1, int32_t c2, int32_t c3, int32_t c4,
                           double *n1, double *n2, double *n3)
{
    const double TWOPI = 6.283185307179586, TINY = 1.0e-35;
    int32_t ctr[4] = {c1, c2, c3, c4};
    double u[4];
    or_philox_uniform(ctr, u);
    double u1 = u[0], u2 = u[1], u3 = u[2], u4 = u[3];
    u1 = fmax(u1, TINY);
    u3 = fmax(u3, TINY);
    double r1 = sqrt(-2.0 * log(u1));
    double th1 = TWOPI * u2;
    double r2 = sqrt(-2.0 * log(u3));
    double th2 = TWOPI * u4;
    *n1 = r1 * cos(th1);
    *n2 = r1 * sin(th1);
    *n3 = r2 * cos(th2);
}

/* ------------------------------------------------------------------ */
/* 3x3 algebra (xmath.f90:130-304).  Column-major: M(i,j) = m[i-1 + 3*(j-1)] */
/* ------------------------------------------------------------------ */
#define M(a, i, j) (a)[((i) - 1) + 3 * ((j) - 1)]

double or_determinant(const double m[9])
{
    return M(m, 1, 1) * (M(m, 2, 2) * M(m, 3, 3) - M(m, 3, 2) * M(m, 2, 3))
         - M(m, 1, 2) * (M(m, 2, 1) * M(m, 3, 3) - M(m, 3, 1) * M(m, 2, 3))
         + M(m, 1, 3) * (M(m, 2, 1) * M(m, 3, 2) - M(m, 3, 1) * M(m, 2, 2));
}

void or_inverse(const double m[9], double out[9])
{
    M(out, 1, 1) = M(m, 2, 2) * M(m, 3, 3) - M(m, 3, 2) * M(m, 2, 3);
    M(out, 1, 2) = M(m, 1, 3) * M(m, 3, 2) - M(m, 1, 2) * M(m, 3, 3);
    M(out, 1, 3) = M(m, 1, 2) * M(m, 2, 3) - M(m, 1, 3) * M(m, 2, 2);
    M(out, 2, 1) = M(m, 2, 3) * M(m, 3, 1) - M(m, 2, 1) * M(m, 3, 3);
    M(out, 2, 2) = M(m, 1, 1) * M(m, 3, 3) - M(m, 1, 3) * M(m, 3, 1);
    M(out, 2, 3) = M(m, 1, 3) * M(m, 2, 1) - M(m, 1, 1) * M(m, 2, 3);
    M(out, 3, 1) = M(m, 2, 1) * M(m, 3, 2) - M(m, 2, 2) * M(m, 3, 1);
    M(out, 3, 2) = M(m, 1, 2) * M(m, 3, 1) - M(m, 1, 1) * M(m, 3, 2);
    M(out, 3, 3) = M(m, 1, 1) * M(m, 2, 2) - M(m, 1, 2) * M(m, 2, 1);
    double det = or_determinant(m);
    for (int k = 0; k < 9; ++k) out[k] = out[k] / det;
}

/* vmdot: a^T . B (xmath.f90:219-229) */
static void vmdot(const double a[3], const double b[9], double out[3])
{
    out[0] = a[0] * M(b, 1, 1) + a[1] * M(b, 2, 1) + a[2] * M(b, 3, 1);
    out[1] = a[0] * M(b, 1, 2) + a[1] * M(b, 2, 2) + a[2] * M(b, 3, 2);
    out[2] = a[0] * M(b, 1, 3) + a[1] * M(b, 2, 3) + a[2] * M(b, 3, 3);
}

/* mvdot: A . b (xmath.f90:231-241) */
static void mvdot(const double a[9], const double b[3], double out[3])
{
    out[0] = M(a, 1, 1) * b[0] + M(a, 1, 2) * b[1] + M(a, 1, 3) * b[2];
    out[1] = M(a, 2, 1) * b[0] + M(a, 2, 2) * b[1] + M(a, 2, 3) * b[2];
    out[2] = M(a, 3, 1) * b[0] + M(a, 3, 2) * b[1] + M(a, 3, 3) * b[2];
}

/* mmdot (xmath.f90:243-260) */
static void mmdot(const double a[9], const double b[9], double out[9])
{
    for (int i = 1; i <= 3; ++i)
        for (int j = 1; j <= 3; ++j)
            M(out, i, j) = M(a, i, 1) * M(b, 1, j) + M(a, i, 2) * M(b, 2, j) + M(a, i, 3) * M(b, 3, j);
}

/* matmul_atba: A^T diag(B) A (xmath.f90:262-279) */
static void matmul_atba(const double A[9], const double B[9], double out[9])
{
    M(out, 1, 1) = M(A, 1, 1) * M(B, 1, 1) * M(A, 1, 1) + M(A, 2, 1) * M(B, 2, 2) * M(A, 2, 1) + M(A, 3, 1) * M(B, 3, 3) * M(A, 3, 1);
    M(out, 1, 2) = M(A, 1, 1) * M(B, 1, 1) * M(A, 1, 2) + M(A, 2, 1) * M(B, 2, 2) * M(A, 2, 2) + M(A, 3, 1) * M(B, 3, 3) * M(A, 3, 2);
    M(out, 1, 3) = M(A, 1, 1) * M(B, 1, 1) * M(A, 1, 3) + M(A, 2, 1) * M(B, 2, 2) * M(A, 2, 3) + M(A, 3, 1) * M(B, 3, 3) * M(A, 3, 3);
    M(out, 2, 1) = M(out, 1, 2);
    M(out, 2, 2) = M(A, 1, 2) * M(B, 1, 1) * M(A, 1, 2) + M(A, 2, 2) * M(B, 2, 2) * M(A, 2, 2) + M(A, 3, 2) * M(B, 3, 3) * M(A, 3, 2);
    M(out, 2, 3) = M(A, 1, 2) * M(B, 1, 1) * M(A, 1, 3) + M(A, 2, 2) * M(B, 2, 2) * M(A, 2, 3) + M(A, 3, 2) * M(B, 3, 3) * M(A, 3, 3);
    M(out, 3, 1) = M(out, 1, 3);
    M(out, 3, 2) = M(out, 2, 3);
    M(out, 3, 3) = M(A, 1, 3) * M(B, 1, 1) * M(A, 1, 3) + M(A, 2, 3) * M(B, 2, 2) * M(A, 2, 3) + M(A, 3, 3) * M(B, 3, 3) * M(A, 3, 3);
}

/* cholesky_solve for 3x3 SPD (xmath.f90:281-304) */
void or_cholesky_solve(const double a[9], const double b[3], double out[3])
{
    double inv_l11 = 1.0 / sqrt(M(a, 1, 1));
    double l21 = M(a, 2, 1) * inv_l11;
    double l31 = M(a, 3, 1) * inv_l11;
    double inv_l22 = 1.0 / sqrt(M(a, 2, 2) - l21 * l21);
    double l32 = (M(a, 3, 2) - l31 * l21) * inv_l22;
    double inv_l33 = 1.0 / sqrt(M(a, 3, 3) - l31 * l31 - l32 * l32);
    double y1 = b[0] * inv_l11;
    double y2 = (b[1] - l21 * y1) * inv_l22;
    double y3 = (b[2] - l31 * y1 - l32 * y2) * inv_l33;
    out[2] = y3 * inv_l33;
    out[1] = (y2 - l32 * out[2]) * inv_l22;
    out[0] = (y1 - l21 * out[1] - l31 * out[2]) * inv_l11;
}

static inline void cross(const double a[3], const double b[3], double out[3])
{
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}

/* x**n for a compile-time integer exponent, evaluated by binary powering
 * (the lowering LLVM's powi uses for the reference's constant exponents). */
static double ipow(double x, int n)
{
    if (n == 0) return 1.0;
    double result = 1.0, base = x;
    int first = 1;
    while (n > 0) {
        if (n & 1) {
            if (first) { result = base; first = 0; }
            else result = result * base;
        }
        n >>= 1;
        if (n) base = base * base;
    }
    return result;
}

/* ------------------------------------------------------------------ */
/* system set-up                                                       */
/* ------------------------------------------------------------------ */

/* set_velocity_to_temperature (system.f90:362-421) */
void or_set_velocity_to_temperature(or_sys *s, double rtemp, int seed, int remove_com)
{
    int n = s->n;
    int temp_int = (int)(rtemp * 1000);
    for (int ix = 1; ix <= n; ++ix) {
        double a, b, c;
        or_philox_normal_3seq(ix, temp_int, seed, 42, &a, &b, &c);
        s->vx[ix - 1] = a; s->vy[ix - 1] = b; s->vz[ix - 1] = c;
    }
    int n_dof = 6 * n - s->nc;
    if (remove_com) {
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int i = 0; i < n; ++i) sx += s->vx[i];
        for (int i = 0; i < n; ++i) sy += s->vy[i];
        for (int i = 0; i < n; ++i) sz += s->vz[i];
        double cx = sx / (double)n, cy = sy / (double)n, cz = sz / (double)n;
        for (int i = 0; i < n; ++i) { s->vx[i] -= cx; s->vy[i] -= cy; s->vz[i] -= cz; }
        n_dof -= 3;
    }
    double ss = 0.0;
    for (int i = 0; i < n; ++i) ss += (s->vx[i] * s->vx[i] + s->vy[i] * s->vy[i]) + s->vz[i] * s->vz[i];
    double factor = sqrt((double)n_dof * rtemp * (double)s->nrigid / ss);
    for (int i = 0; i < n; ++i) { s->vx[i] *= factor; s->vy[i] *= factor; s->vz[i] *= factor; }
}

/* defaults applied by initialize_system (system.f90:566-718) to arrays left
 * at their set_num_atoms sentinel (-1) */
void or_initialize_defaults(or_sys *s, int remove_com)
{
    s->dof = 6 * s->n - s->nc;
    if (remove_com) s->dof -= 3;
    for (int i = 0; i < s->n; ++i) {
        if (s->charge[i] == -1 || s->charge[i] < 0) s->charge[i] = 0.0;
        if (s->sig_x[i] == -1 || s->sig_x[i] < 0) { s->sig_x[i] = 1.0; s->sig_y[i] = 1.0; s->sig_z[i] = 1.0; }
        if (s->eps_x[i] == -1 || s->eps_x[i] < 0) { s->eps_x[i] = 1.0; s->eps_y[i] = 1.0; s->eps_z[i] = 1.0; }
        int set_inertia_from_shape = 0;
        if (s->mass[i] == -1 || s->mass[i] < 0) {
            s->mass[i] = 1.0; s->inv_mass[i] = 1.0;
            set_inertia_from_shape = 1;
        } else if (s->mass[i] == 0) {
            s->inv_mass[i] = 0.0;
            s->ixx[i] = s->iyy[i] = s->izz[i] = 0.0;
            s->ixx_inv[i] = s->iyy_inv[i] = s->izz_inv[i] = 0.0;
            continue;
        } else {
            s->inv_mass[i] = 1.0 / s->mass[i];
            set_inertia_from_shape = 1;
        }
        if (set_inertia_from_shape) {
            if (s->ixx[i] < 0 || s->iyy[i] < 0 || s->izz[i] < 0) {
                s->ixx[i] = (s->sig_y[i] * s->sig_y[i] + s->sig_z[i] * s->sig_z[i]) / 20.0;
                s->iyy[i] = (s->sig_x[i] * s->sig_x[i] + s->sig_z[i] * s->sig_z[i]) / 20.0;
                s->izz[i] = (s->sig_x[i] * s->sig_x[i] + s->sig_y[i] * s->sig_y[i]) / 20.0;
                s->ixx_inv[i] = 1.0 / s->ixx[i];
                s->iyy_inv[i] = 1.0 / s->iyy[i];
                s->izz_inv[i] = 1.0 / s->izz[i];
            } else if (s->ixx[i] > 0 && s->iyy[i] > 0 && s->izz[i] > 0) {
                s->ixx_inv[i] = 1.0 / s->ixx[i];
                s->iyy_inv[i] = 1.0 / s->iyy[i];
                s->izz_inv[i] = 1.0 / s->izz[i];
            } else {
                s->ixx_inv[i] = s->iyy_inv[i] = s->izz_inv[i] = 0.0;
            }
        }
    }
    or_initialize_forces(s);
}

/* initialize_forces (system.f90:423-436) */
void or_initialize_forces(or_sys *s)
{
    for (int i = 0; i < s->n; ++i) {
        s->fx[i] = s->fy[i] = s->fz[i] = 0.0;
        s->tx[i] = s->ty[i] = s->tz[i] = 0.0;
    }
    s->pot_eng = s->kin_eng = s->kin_rot = s->temp = 0.0;
    s->pressure = s->virial = s->virial_bonded = 0.0;
}

/* calculate_thermo (system.f90:725-741) */
void or_calculate_thermo(or_sys *s)
{
    double kin_tot = s->kin_eng + s->kin_rot;
    s->temp = 2.0 * kin_tot / (double)s->dof;
    double w = 0.0;
    for (int i = 0; i < s->n; ++i) w += s->rx[i] * s->fx[i] + s->ry[i] * s->fy[i] + s->rz[i] * s->fz[i];
    if (s->nc > 0 && s->virial_bonded >= 0.0) w = w + s->virial_bonded;
    s->virial = w;
    double volume = s->box[0] * s->box[1] * s->box[2];
    s->pressure = (s->density * s->nrigid * s->temp) + (s->virial / (3.0 * volume));
}

/* generate_fcc_lattice (system.f90:438-523) with set_density (:167-183) */
void or_generate_fcc_lattice(or_sys *s, int unit_cells, double dens)
{
    int num_atoms = 4 * unit_cells * unit_cells * unit_cells;
    s->density = dens;
    double vol = (double)(s->n / s->nrigid) / dens;
    double box_len = pow(vol, 1.0 / 3.0);
    s->box[0] = s->box[1] = s->box[2] = box_len;
    double cell = s->box[0] / (double)unit_cells;
    double half_cell = cell / 2.0;
    double rroot3 = 1.0 / sqrt(3.0);
    double den = sqrt(1.0 - rroot3 * rroot3);
    double sp = sqrt((1.0 + rroot3) / 2.0), sm = sqrt((1.0 - rroot3) / 2.0);
    s->rx[0] = 0.0; s->ry[0] = 0.0; s->rz[0] = 0.0;
    s->qw[0] = sp; s->qx[0] = sm * (rroot3 / den); s->qy[0] = sm * (-rroot3 / den); s->qz[0] = 0.0;
    s->rx[1] = half_cell; s->ry[1] = half_cell; s->rz[1] = 0.0;
    s->qw[1] = sm; s->qx[1] = sp * (-rroot3 / den); s->qy[1] = sp * (-rroot3 / den); s->qz[1] = 0.0;
    s->rx[2] = 0.0; s->ry[2] = half_cell; s->rz[2] = half_cell;
    s->qw[2] = sm; s->qx[2] = sp * (rroot3 / den); s->qy[2] = sp * (rroot3 / den); s->qz[2] = 0.0;
    s->rx[3] = half_cell; s->ry[3] = 0.0; s->rz[3] = half_cell;
    s->qw[3] = sp; s->qx[3] = sm * (-rroot3 / den); s->qy[3] = sm * (rroot3 / den); s->qz[3] = 0.0;
    int mtemp = 0;
    for (int I = 1; I <= unit_cells; ++I)
        for (int J = 1; J <= unit_cells; ++J)
            for (int K = 1; K <= unit_cells; ++K) {
                for (int iref = 0; iref < 4; ++iref) {
                    s->rx[iref + mtemp] = s->rx[iref] + cell * (double)(I - 1);
                    s->ry[iref + mtemp] = s->ry[iref] + cell * (double)(J - 1);
                    s->rz[iref + mtemp] = s->rz[iref] + cell * (double)(K - 1);
                    s->qw[iref + mtemp] = s->qw[iref];
                    s->qx[iref + mtemp] = s->qx[iref];
                    s->qy[iref + mtemp] = s->qy[iref];
                    s->qz[iref + mtemp] = s->qz[iref];
                }
                mtemp += 4;
            }
    for (int i = 0; i < num_atoms; ++i) {
        s->rx[i] = s->rx[i] - s->box[0] / 2.0;
        s->ry[i] = s->ry[i] - s->box[1] / 2.0;
        s->rz[i] = s->rz[i] - s->box[2] / 2.0;
    }
}

/* ------------------------------------------------------------------ */
/* neighbour list (system.f90:927-1039)                                */
/* ------------------------------------------------------------------ */
int or_nlist_build(or_sys *s, or_nlist *nl)
{
    double rskin = nl->cutoff + nl->skin;
    nl->r_skin_sq = rskin * rskin;
    long K = 0;
    int n = s->n;
    for (int i = 0; i < n - 1; ++i) {
        nl->offset[i] = K;
        for (int j = i + 1; j < n; ++j) {
            double d0 = s->rx[i] - s->rx[j], d1 = s->ry[i] - s->ry[j], d2 = s->rz[i] - s->rz[j];
            if (s->periodic) {
                d0 = mimg(d0, s->box[0]); d1 = mimg(d1, s->box[1]); d2 = mimg(d2, s->box[2]);
            }
            double rsq = d0 * d0 + d1 * d1 + d2 * d2;
            if (rsq < nl->r_skin_sq) {
                if (K >= nl->cap) return -1;
                nl->nbr[K++] = j;
            }
        }
    }
    if (n >= 1) nl->offset[n - 1] = K;
    nl->offset[n] = K;
    nl->nnz = K;
    memcpy(nl->x_last, s->rx, sizeof(double) * n);
    memcpy(nl->y_last, s->ry, sizeof(double) * n);
    memcpy(nl->z_last, s->rz, sizeof(double) * n);
    return 0;
}

int or_nlist_needs_rebuild(const or_sys *s, const or_nlist *nl)
{
    double max_disp = 0.0;
    for (int i = 0; i < s->n; ++i) {
        double dx = s->rx[i] - nl->x_last[i];
        double dy = s->ry[i] - nl->y_last[i];
        double dz = s->rz[i] - nl->z_last[i];
        if (s->periodic) {
            dx = mimg(dx, s->box[0]); dy = mimg(dy, s->box[1]); dz = mimg(dz, s->box[2]);
        }
        double r = dx * dx + dy * dy + dz * dz;
        max_disp = fmax(max_disp, r);
    }
    return (4.0 * max_disp > nl->r_skin_sq) ? 1 : 0;
}

/* ------------------------------------------------------------------ */
/* isotropic pair kernels                                              */
/* ------------------------------------------------------------------ */
static inline void pair_rij(const or_sys *s, int i, int j, double r[3])
{
    r[0] = s->rx[i] - s->rx[j];
    r[1] = s->ry[i] - s->ry[j];
    r[2] = s->rz[i] - s->rz[j];
    if (s->periodic) {
        r[0] = mimg(r[0], s->box[0]);
        r[1] = mimg(r[1], s->box[1]);
        r[2] = mimg(r[2], s->box[2]);
    }
}

/* lj_calculate_forces (force.f90:6-49) */
void or_lj(const or_sys *s, int i, int j, double sigma, double epsilon, double cutoff,
           double *pe, double f[3])
{
    double r[3];
    pair_rij(s, i, j, r);
    double rij_sq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    double rcut_sq = cutoff * cutoff;
    int cmask = rij_sq < rcut_sq;
    double sr2 = (sigma * sigma) / rij_sq, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
    double sc2 = (sigma * sigma) / rcut_sq, sc6 = sc2 * sc2 * sc2, sc12 = sc6 * sc6;
    double pot = 4.0 * epsilon * (sr12 - sr6);
    double pot_cut = 4.0 * epsilon * (sc12 - sc6);
    double coeff = 24.0 * epsilon * (2.0 * sr12 - sr6);
    *pe = cmask ? (pot - pot_cut) : 0.0;
    for (int k = 0; k < 3; ++k) f[k] = cmask ? (coeff * r[k] / rij_sq) : 0.0;
}

/* lj12_calculate_forces (force.f90:51-93) */
void or_lj12(const or_sys *s, int i, int j, double sigma, double epsilon, double cutoff,
             double *pe, double f[3])
{
    double r[3];
    pair_rij(s, i, j, r);
    *pe = 0.0; f[0] = f[1] = f[2] = 0.0;
    double rij_sq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    double rcut_sq = cutoff * cutoff;
    if (rij_sq < rcut_sq) {
        double sr2 = (sigma * sigma) / rij_sq, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
        double sc2 = (sigma * sigma) / rcut_sq, sc6 = sc2 * sc2 * sc2, sc12 = sc6 * sc6;
        double pot = epsilon * (sr12 - 2.0 * sr6);
        double pot_cut = epsilon * (sc12 - 2.0 * sc6);
        double coeff = 12.0 * epsilon * (sr12 - sr6);
        *pe = pot - pot_cut;
        for (int k = 0; k < 3; ++k) f[k] = coeff * r[k] / rij_sq;
    }
}

/* ljw_calculate_forces (force.f90:95-146) */
void or_ljw(const or_sys *s, int i, int j, double sigma, double epsilon, double cutoff,
            double wf, double *pe, double f[3])
{
    double r[3];
    pair_rij(s, i, j, r);
    *pe = 0.0; f[0] = f[1] = f[2] = 0.0;
    double rij_sq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    double rij_mag = sqrt(rij_sq);
    if (rij_mag < (cutoff + wf)) {
        double sr;
        if (rij_mag < sigma) sr = sigma / rij_mag;
        else if ((rij_mag >= sigma) && (rij_mag < (sigma + wf))) sr = 1.0;
        else sr = sigma / (rij_mag - wf);
        double sr2 = sr * sr, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
        double sc = sigma / cutoff, sc2 = sc * sc, sc6 = sc2 * sc2 * sc2, sc12 = sc6 * sc6;
        double pot = epsilon * (sr12 - 2.0 * sr6);
        double pot_cut = epsilon * (sc12 - 2.0 * sc6);
        double coeff = 12.0 * epsilon * (sr12 - sr6) * (sr / sigma);
        *pe = pot - pot_cut;
        for (int k = 0; k < 3; ++k) f[k] = coeff * r[k] / rij_mag;
    }
}

/* wca_calculate_forces (force.f90:155-190) */
void or_wca(const or_sys *s, int i, int j, double sigma, double epsilon,
            double *pe, double f[3])
{
    double r[3];
    pair_rij(s, i, j, r);
    *pe = 0.0; f[0] = f[1] = f[2] = 0.0;
    double rij_sq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    double sigma_sq = sigma * sigma;
    if (rij_sq <= sigma_sq) {
        double sr2 = sigma_sq / rij_sq, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
        *pe = epsilon * (sr12 - 2.0 * sr6) + epsilon;
        double coeff = 12.0 * epsilon * (sr12 - sr6);
        for (int k = 0; k < 3; ++k) f[k] = coeff * r[k] / rij_sq;
    }
}

/* morse_calculate_forces (force.f90:199-233) */
void or_morse(const or_sys *s, int i, int j, double dissoc, double width, double rmin,
              double cutoff, double *pe, double f[3])
{
    double r[3];
    pair_rij(s, i, j, r);
    *pe = 0.0; f[0] = f[1] = f[2] = 0.0;
    double rij_sq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    double rij_mag = sqrt(rij_sq);
    if (rij_mag < cutoff) {
        double e = exp(-width * (rij_mag - rmin));
        double ec = exp(-width * (cutoff - rmin));
        double pot = dissoc * ((1.0 - e) * (1.0 - e) - 1.0);
        double pot_cut = dissoc * ((1.0 - ec) * (1.0 - ec) - 1.0);
        *pe = pot - pot_cut;
        for (int k = 0; k < 3; ++k) f[k] = -2.0 * width * dissoc * (1.0 - e) * e * r[k] / rij_mag;
    }
}

/* dh_calculate_forces (force.f90:2589-2617) */
void or_dh(const or_sys *s, int i, int j, double cutoff, double db_len, double dielectric,
           double *pe, double f[3])
{
    double r[3];
    pair_rij(s, i, j, r);
    double rij_sq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    double rij_mag = sqrt(rij_sq);
    int cmask = rij_mag < cutoff;
    double p = (s->charge[i] * s->charge[j] * exp(-rij_mag / db_len)) / (4.0 * PI_REF * dielectric * rij_mag);
    double fac = p * (1.0 + rij_mag / db_len);
    *pe = cmask ? p : 0.0;
    for (int k = 0; k < 3; ++k) f[k] = cmask ? (fac * r[k] / rij_sq) : 0.0;
}

/* coulombic_calculate_forces (force.f90:2627-2665): q_i q_j / (4 pi eps r),
   switched off over the last 2 units before the cutoff; the switched force
   subtracts sdx * PE * factor (PE already switched) from every component */
void or_coulomb(const or_sys *s, int i, int j, double cutoff, double dielectric, double *pe, double f[3])
{
    *pe = 0.0;
    f[0] = f[1] = f[2] = 0.0;
    double r[3];
    r[0] = s->rx[i] - s->rx[j];
    r[1] = s->ry[i] - s->ry[j];
    r[2] = s->rz[i] - s->rz[j];
    if (s->periodic)
        for (int k = 0; k < 3; ++k) r[k] = mimg(r[k], s->box[k]);
    double rij_sq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    double rij_mag = sqrt(rij_sq);
    double sw = cutoff - 2.0;
    double factor = 1.0 / (cutoff - sw);
    double xx = (rij_mag - sw) * factor;
    double sx = 1.0 - 6.0 * xx * xx * xx * xx * xx + 15.0 * xx * xx * xx * xx - 10.0 * xx * xx * xx;
    double sdx = -30.0 * xx * xx * xx * xx + 60.0 * xx * xx * xx - 30.0 * xx * xx;
    if (rij_mag < cutoff) {
        double p = s->charge[i] * s->charge[j] / (4.0 * PI_REF * dielectric * rij_mag);
        for (int k = 0; k < 3; ++k) f[k] = p * r[k] / rij_sq;
        if (rij_mag >= sw) {
            p = sx * p;
            for (int k = 0; k < 3; ++k) f[k] = sx * f[k] - sdx * p * factor;
        }
        *pe = p;
    }
}

/* ------------------------------------------------------------------ */
/* composite rigid bodies (system.f90:743-881)                          */
/* ------------------------------------------------------------------ */
static void inertia_from_shape(or_sys *s, int i)
{
    s->ixx[i] = (pow(s->sig_y[i], 2.0) + pow(s->sig_z[i], 2.0)) / 20.0;
    s->iyy[i] = (pow(s->sig_x[i], 2.0) + pow(s->sig_z[i], 2.0)) / 20.0;
    s->izz[i] = (pow(s->sig_x[i], 2.0) + pow(s->sig_y[i], 2.0)) / 20.0;
}

static void inertia_add(or_sys *s, int from, int to, double xo, double yo, double zo)
{
    s->ixx[from] = s->ixx[from] + s->ixx[to] + (pow(yo, 2.0) + pow(zo, 2.0));
    s->iyy[from] = s->iyy[from] + s->iyy[to] + (pow(xo, 2.0) + pow(zo, 2.0));
    s->izz[from] = s->izz[from] + s->izz[to] + (pow(xo, 2.0) + pow(yo, 2.0));
}

/* copy_transform (system.f90:743-842): places `to` at `from` plus the offset
   rotated by from's orientation and gives it from's quaternion; set_inertia
   -1 (absent), 0 or 1.  Returns 0, or 1/2/3 for the three stops (second
   particle without shape, first particle without shape, second particle
   with mass) */
int or_copy_transform(or_sys *s, int from, int to, double xo, double yo, double zo, int set_inertia)
{
    double w = s->qw[from], x = s->qx[from], y = s->qy[from], z = s->qz[from];
    double rot[3][3];
    rot[0][0] = w * w + x * x - y * y - z * z;
    rot[0][1] = 2.0 * (x * y - w * z);
    rot[0][2] = 2.0 * (x * z + w * y);
    rot[1][0] = 2.0 * (x * y + w * z);
    rot[1][1] = w * w - x * x + y * y - z * z;
    rot[1][2] = 2.0 * (y * z - w * x);
    rot[2][0] = 2.0 * (x * z - w * y);
    rot[2][1] = 2.0 * (y * z + w * x);
    rot[2][2] = w * w - x * x - y * y + z * z;
    double ro[3];
    for (int k = 0; k < 3; ++k) ro[k] = rot[k][0] * xo + rot[k][1] * yo + rot[k][2] * zo;
    s->rx[to] = s->rx[from] + ro[0];
    s->ry[to] = s->ry[from] + ro[1];
    s->rz[to] = s->rz[from] + ro[2];
    s->qw[to] = s->qw[from];
    s->qx[to] = s->qx[from];
    s->qy[to] = s->qy[from];
    s->qz[to] = s->qz[from];
    if (set_inertia != 1) return 0;
    if (!(s->mass[to] <= 0.0)) return 3;
    int to_has_i = s->ixx[to] > 0.0 && s->iyy[to] > 0.0 && s->izz[to] > 0.0;
    int to_has_shape = s->sig_x[to] > 0.0 && s->sig_y[to] > 0.0 && s->sig_z[to] > 0.0;
    if (s->ixx[from] > 0.0 && s->iyy[from] > 0.0 && s->izz[from] > 0.0) {
    } else if (s->sig_x[from] > 0.0 && s->sig_y[from] > 0.0 && s->sig_z[from] > 0.0) {
        inertia_from_shape(s, from);
    } else {
        return 2;
    }
    if (to_has_i) {
    } else if (to_has_shape) {
        inertia_from_shape(s, to);
    } else {
        return 1;
    }
    inertia_add(s, from, to, xo, yo, zo);
    return 0;
}

/* transfer_forces (system.f90:844-881): from's force and torque, plus the
   torque of its force about `to`, move to `to` */
void or_transfer_forces(or_sys *s, int from, int to)
{
    double r[3] = {s->rx[from] - s->rx[to], s->ry[from] - s->ry[to], s->rz[from] - s->rz[to]};
    if (s->periodic)
        for (int k = 0; k < 3; ++k) r[k] = mimg(r[k], s->box[k]);
    double f[3] = {s->fx[from], s->fy[from], s->fz[from]};
    double t[3];
    t[0] = r[1] * f[2] - r[2] * f[1];
    t[1] = r[2] * f[0] - r[0] * f[2];
    t[2] = r[0] * f[1] - r[1] * f[0];
    s->fx[to] = s->fx[to] + s->fx[from];
    s->fy[to] = s->fy[to] + s->fy[from];
    s->fz[to] = s->fz[to] + s->fz[from];
    s->tx[to] = s->tx[to] + s->tx[from] + t[0];
    s->ty[to] = s->ty[to] + s->ty[from] + t[1];
    s->tz[to] = s->tz[to] + s->tz[from] + t[2];
    s->fx[from] = 0.0; s->fy[from] = 0.0; s->fz[from] = 0.0;
    s->tx[from] = 0.0; s->ty[from] = 0.0; s->tz[from] = 0.0;
}

/* ------------------------------------------------------------------ */
/* point dipoles with a reaction field (module dipole, force.f90:2669-2767) */
/* ------------------------------------------------------------------ */
/* the body z axis scaled by the particle's charge (force.f90:2709-2718) */
static void dip_axis(const or_sys *s, int i, double u[3])
{
    double w = s->qw[i], x = s->qx[i], y = s->qy[i], z = s->qz[i];
    u[0] = 2.0 * (x * z + w * y);
    u[1] = 2.0 * (y * z - w * x);
    u[2] = w * w - x * x - y * y + z * z;
    for (int k = 0; k < 3; ++k) u[k] = u[k] * s->charge[i];
}

/* dipole_calculate_forces (force.f90:2687-2751); dip and rfl are the
   module's DIPX..DIPZ and RFLX..RFLZ (3 arrays of n each).  A pair within
   the cutoff sets particle i's dipole and adds particle j's scaled dipole to
   the reaction field of EVERY particle (the whole-array update :2745-2747). */
void or_dipole(const or_sys *s, int i, int j, double cutoff, double dielectric, double *pe, double f[3],
               double ti[3], double tj[3], double *dip, double *rfl)
{
    int n = s->n;
    *pe = 0.0;
    for (int k = 0; k < 3; ++k) f[k] = ti[k] = tj[k] = 0.0;
    double dip_coeff = (2.0 * (dielectric - 1.0)) / (2.0 * dielectric + 1.0);
    dip_coeff = dip_coeff / (cutoff * cutoff * cutoff);
    double r[3];
    r[0] = s->rx[i] - s->rx[j];
    r[1] = s->ry[i] - s->ry[j];
    r[2] = s->rz[i] - s->rz[j];
    if (s->periodic)
        for (int k = 0; k < 3; ++k) r[k] = mimg(r[k], s->box[k]);
    double rij_sq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    double rij_mag = sqrt(rij_sq);
    double u1[3], u2[3];
    dip_axis(s, i, u1);
    dip_axis(s, j, u2);
    if (!(rij_mag < cutoff)) return;
    double ci = u1[0] * r[0] + u1[1] * r[1] + u1[2] * r[2];
    double cj = u2[0] * r[0] + u2[1] * r[1] + u2[2] * r[2];
    double cij = u1[0] * u2[0] + u1[1] * u2[1] + u1[2] * u2[2];
    double t1[3], t2[3], t3[3];
    for (int k = 0; k < 3; ++k) {
        t1[k] = (cij - (5.0 * ci * cj / rij_sq)) * r[k] + cj * u1[k] + ci * u2[k];
        t2[k] = 3.0 * cj * (r[k] / rij_sq) - u2[k];
        t3[k] = 3.0 * ci * (r[k] / rij_sq) - u1[k];
    }
    double rij_cube = rij_sq * rij_mag;
    double rij_pow5 = rij_cube * rij_sq;
    *pe = (cij / rij_cube) - ((3.0 * ci * cj) / rij_pow5);
    double c1[3], c2[3];
    cross(u1, t2, c1);
    cross(u2, t3, c2);
    for (int k = 0; k < 3; ++k) {
        f[k] = (3.0 / rij_pow5) * t1[k];
        ti[k] = (1.0 / rij_cube) * c1[k];
        tj[k] = (1.0 / rij_cube) * c2[k];
    }
    for (int k = 0; k < 3; ++k) dip[k * n + i] = u1[k];
    for (int k = 0; k < 3; ++k) {
        double add = u2[k] * dip_coeff;
        for (int q = 0; q < n; ++q) rfl[k * n + q] = rfl[k * n + q] + add;
    }
}

/* eval_reaction_field (force.f90:2753-2766) */
void or_eval_reaction_field(or_sys *s, const double *dip, const double *rfl)
{
    int n = s->n;
    for (int i = 0; i < n; ++i) {
        if (s->charge[i] != 0.0) {
            double dx = dip[i], dy = dip[n + i], dz = dip[2 * n + i];
            double rx = rfl[i], ry = rfl[n + i], rz = rfl[2 * n + i];
            s->pot_eng = s->pot_eng + 0.5 * (dx * rx + dy * ry + dz * rz);
            s->tx[i] = s->tx[i] - (dy * rz - dz * ry);
            s->ty[i] = s->ty[i] - (dz * rx - dx * rz);
            s->tz[i] = s->tz[i] - (dx * ry - dy * rx);
        }
    }
}

/* the caller's pair loop (README.md:86-103) over dipole_calculate_forces */
void or_dipole_loop(or_sys *s, const or_nlist *nl, double cutoff, double dielectric, double *dip, double *rfl)
{
    for (int i = 0; i < s->n - 1; ++i) {
        for (long k = nl->offset[i]; k < nl->offset[i + 1]; ++k) {
            int j = nl->nbr[k];
            double pe, f[3], ti[3], tj[3];
            or_dipole(s, i, j, cutoff, dielectric, &pe, f, ti, tj, dip, rfl);
            s->pot_eng = s->pot_eng + pe;
            s->fx[i] = s->fx[i] + f[0]; s->fy[i] = s->fy[i] + f[1]; s->fz[i] = s->fz[i] + f[2];
            s->fx[j] = s->fx[j] - f[0]; s->fy[j] = s->fy[j] - f[1]; s->fz[j] = s->fz[j] - f[2];
            s->tx[i] = s->tx[i] + ti[0]; s->ty[i] = s->ty[i] + ti[1]; s->tz[i] = s->tz[i] + ti[2];
            s->tx[j] = s->tx[j] + tj[0]; s->ty[j] = s->ty[j] + tj[1]; s->tz[j] = s->tz[j] + tj[2];
        }
    }
}

/* ------------------------------------------------------------------ */
/* Gay-Berne (force.f90:237-533)                                       */
/* ------------------------------------------------------------------ */
typedef struct { double ru1, ru2, uu; } gb_par;
typedef struct { double u1[3], u2[3], r[3], rsq; } gb_pr;

static double g_func(double chie, const gb_par *p, const gb_pr *q)
{
    double t1 = (p->ru1 + p->ru2) * (p->ru1 + p->ru2) / (1 + chie * p->uu);
    double t2 = (p->ru1 - p->ru2) * (p->ru1 - p->ru2) / (1 - chie * p->uu);
    return 1 - (chie / 2.0 / q->rsq) * (t1 + t2);
}

static void dG_dr(double chie, const gb_par *p, const gb_pr *q, double res[3])
{
    double a = (p->ru1 + p->ru2) / (1 + chie * p->uu);
    double b = (p->ru1 - p->ru2) / (1 - chie * p->uu);
    double g = g_func(chie, p, q);
    for (int k = 0; k < 3; ++k) {
        double t1 = a * (q->u1[k] + q->u2[k]);
        double t2 = b * (q->u1[k] - q->u2[k]);
        res[k] = (-chie / q->rsq) * (t1 + t2) + (2 * q->r[k] / q->rsq) * (1 - g);
    }
}

static void dG_du1(double chie, const gb_par *p, const gb_pr *q, double res[3])
{
    double t1 = (p->ru1 + p->ru2) / (1 + chie * p->uu);
    double t2 = (p->ru1 - p->ru2) / (1 - chie * p->uu);
    for (int k = 0; k < 3; ++k)
        res[k] = (-chie * q->r[k] / q->rsq) * (t1 + t2) + ((chie * chie * q->u2[k]) / (2 * q->rsq)) * (t1 * t1 - t2 * t2);
}

static void dG_du2(double chie, const gb_par *p, const gb_pr *q, double res[3])
{
    double t1 = (p->ru1 + p->ru2) / (1 + chie * p->uu);
    double t2 = (p->ru1 - p->ru2) / (1 - chie * p->uu);
    for (int k = 0; k < 3; ++k)
        res[k] = (-chie * q->r[k] / q->rsq) * (t1 - t2) + ((chie * chie * q->u1[k]) / (2 * q->rsq)) * (t1 * t1 - t2 * t2);
}

static inline void quat_u(const or_sys *s, int i, double u[3])
{
    double w = s->qw[i], x = s->qx[i], y = s->qy[i], z = s->qz[i];
    u[0] = 2.0 * (x * z + w * y);
    u[1] = 2.0 * (y * z - w * x);
    u[2] = w * w - x * x - y * y + z * z;
}

/* gb_ (variant 0, force.f90:308-428), gbr_ (1, :430-533), gbw_ (2, welded:
   :535-663), gbc_ (3, chiral: :665-843), gbwc_ (4, welded and chiral:
   :845-1030).  The chiral forms regroup the plain terms (the energy factors
   move into dR and the srx/srd factors out, :694-708) before adding the
   chirality-weighted pseudoscalar terms (:712-759). */
void or_gbx(const or_sys *s, int i, int j, int variant, double cutoff, double wf, double chirality, int mu, int nu,
            double *pe, double f[3], double ti[3], double tj[3])
{
    gb_par par;
    gb_pr pr;
    const int welded = (variant == 2 || variant == 4), chiral = (variant == 3 || variant == 4);
    double pot = 0.0, FIJ[3] = {0, 0, 0}, T1[3] = {0, 0, 0}, T2[3] = {0, 0, 0};
    const double sigma_0 = 1.0;
    pair_rij(s, i, j, pr.r);
    pr.rsq = pr.r[0] * pr.r[0] + pr.r[1] * pr.r[1] + pr.r[2] * pr.r[2];
    double rij_mag = sqrt(pr.rsq);
    double rhat[3] = {pr.r[0] / rij_mag, pr.r[1] / rij_mag, pr.r[2] / rij_mag};
    quat_u(s, i, pr.u1);
    quat_u(s, j, pr.u2);
    double lpar = s->sig_z[i], lprp = s->sig_x[i];
    double epar = s->eps_z[i], eprp = s->eps_x[i];
    double muinv = 1.0 / mu;
    par.ru1 = pr.r[0] * pr.u1[0] + pr.r[1] * pr.u1[1] + pr.r[2] * pr.u1[2];
    par.ru2 = pr.r[0] * pr.u2[0] + pr.r[1] * pr.u2[1] + pr.r[2] * pr.u2[2];
    par.uu = pr.u1[0] * pr.u2[0] + pr.u1[1] * pr.u2[1] + pr.u1[2] * pr.u2[2];
    double chi = (lpar * lpar - lprp * lprp) / (lpar * lpar + lprp * lprp);
    double ep = (mu == 1) ? eprp : pow(eprp, muinv);
    double ea = (mu == 1) ? epar : pow(epar, muinv);
    double xhi = (ep - ea) / (ep + ea);
    double gchi = g_func(chi, &par, &pr);
    double gxhi = g_func(xhi, &par, &pr);
    double eps1 = 1.0 / sqrt(1 - (chi * par.uu) * (chi * par.uu));
    double eps2 = gxhi / s->eps_z[i];
    double sigma = lprp / sqrt(gchi);
    double sigma_0_inv = 1.0 / sigma_0;
    double g32 = sigma / gchi;
    double switch_ = 0, factor = 0, sx = 0, sdx = 0, rlim;
    if (variant != 1) {
        rlim = welded ? (sigma + cutoff + wf - sigma_0) : (sigma + cutoff - sigma_0);
        switch_ = 0.9 * rlim;
        factor = 1.0 / (rlim - switch_);
        double xx = (rij_mag - switch_) * factor;
        sx = 1.0 - 6.0 * xx * xx * xx * xx * xx + 15.0 * xx * xx * xx * xx - 10.0 * xx * xx * xx;
        sdx = -30.0 * xx * xx * xx * xx + 60.0 * xx * xx * xx - 30.0 * xx * xx;
    } else {
        rlim = sigma;
    }
    int inside = (variant == 1) ? (rij_mag <= rlim) : (rij_mag < rlim);
    if (inside) {
        double sr;
        if (!welded) sr = sigma_0 / (rij_mag - sigma + sigma_0);
        else if (rij_mag < sigma) sr = sigma_0 / (rij_mag - sigma + sigma_0);
        else if (rij_mag >= sigma && rij_mag <= sigma + wf) sr = 1.0;
        else sr = sigma_0 / (rij_mag - sigma + sigma_0 - wf);
        double sr2 = sr * sr, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
        double srx = (sr12 - 2.0 * sr6);
        if (variant == 1) srx = srx + 1.0;
        double srd = 12.0 * (sr6 - sr12) * sr;
        double e1n = ipow(eps1, nu), e2m = ipow(eps2, mu);
        double e2m1 = ipow(eps2, mu - 1), e1n2 = ipow(eps1, nu + 2);
        double dgr_c[3], dgr_x[3], dgu1_x[3], dgu1_c[3], dgu2_x[3], dgu2_c[3];
        dG_dr(chi, &par, &pr, dgr_c);
        dG_dr(xhi, &par, &pr, dgr_x);
        dG_du1(xhi, &par, &pr, dgu1_x);
        dG_du1(chi, &par, &pr, dgu1_c);
        dG_du2(xhi, &par, &pr, dgu2_x);
        dG_du2(chi, &par, &pr, dgu2_c);
        double TI[3], TJ[3];
        double dR_dr[3], de2_dr[3], de1_du1[3], de2_du1[3], dR_du1[3], de1_du2[3], de2_du2[3], dR_du2[3];
        for (int k = 0; k < 3; ++k) {
            de2_dr[k] = e1n * (mu * e2m1) * dgr_x[k];
            de1_du1[k] = e2m * (nu * e1n2) * (chi * chi * par.uu * pr.u2[k]);
            de2_du1[k] = e1n * (mu * e2m1) * dgu1_x[k];
            de1_du2[k] = e2m * (nu * e1n2) * (chi * chi * par.uu * pr.u1[k]);
            de2_du2[k] = e1n * (mu * e2m1) * dgu2_x[k];
            if (!chiral) {
                dR_dr[k] = sigma_0_inv * (rhat[k] + 0.5 * dgr_c[k] * g32);
                dR_du1[k] = srd * 0.5 * sigma_0_inv * dgu1_c[k] * g32;
                dR_du2[k] = srd * 0.5 * sigma_0_inv * dgu2_c[k] * g32;
                FIJ[k] = (-1.0) * (e1n * e2m * srd * dR_dr[k] + de2_dr[k] * srx);
                TI[k] = (-1.0) * (srx * (de1_du1[k] + de2_du1[k]) + (e1n * e2m * dR_du1[k]));
                TJ[k] = (-1.0) * (srx * (de1_du2[k] + de2_du2[k]) + (e1n * e2m * dR_du2[k]));
            } else {
                dR_dr[k] = e1n * e2m * sigma_0_inv * (rhat[k] + 0.5 * dgr_c[k] * g32);
                dR_du1[k] = e1n * e2m * 0.5 * sigma_0_inv * dgu1_c[k] * g32;
                dR_du2[k] = e1n * e2m * 0.5 * sigma_0_inv * dgu2_c[k] * g32;
                FIJ[k] = (-1.0) * (srx * de2_dr[k] + srd * dR_dr[k]);
                TI[k] = (-1.0) * (srx * (de1_du1[k] + de2_du1[k]) + srd * dR_du1[k]);
                TJ[k] = (-1.0) * (srx * (de1_du2[k] + de2_du2[k]) + srd * dR_du2[k]);
            }
        }
        pot = e1n * e2m * srx;
        cross(pr.u1, TI, T1);
        cross(pr.u2, TJ, T2);
        if (chiral) {
            double srxc, srdc;
            if (welded) {
                srxc = (sr12 - (12.0 / 7.0) * sr6 * sr);
                srdc = 12.0 * (sr6 * sr2 - sr12 * sr);
            } else {
                srxc = sr6 * sr;
                srdc = -7.0 * sr6 * sr2;
            }
            double u1xu2[3], u2xrij[3], rijxu1[3];
            cross(pr.u1, pr.u2, u1xu2);
            cross(pr.u2, rhat, u2xrij);
            cross(rhat, pr.u1, rijxu1);
            double uur = u1xu2[0] * rhat[0] + u1xu2[1] * rhat[1] + u1xu2[2] * rhat[2];
            double uu = par.uu;
            double FC[3], TIc[3], TJc[3], T1c[3], T2c[3];
            for (int k = 0; k < 3; ++k) {
                double dRr = dR_dr[k] * srdc * uur * uu;
                double de2r = de2_dr[k] * srxc * uur * uu;
                double duur_r = e1n * e2m * srxc * uu * (u1xu2[k] - uur * rhat[k]) / rij_mag;
                double a1 = de1_du1[k] * srxc * uur * uu, b1 = de2_du1[k] * srxc * uur * uu;
                double c1 = dR_du1[k] * srdc * uur * uu;
                double duur1 = e1n * e2m * srxc * uu * u2xrij[k], duu1 = e1n * e2m * srxc * uur * pr.u2[k];
                double a2 = de1_du2[k] * srxc * uur * uu, b2 = de2_du2[k] * srxc * uur * uu;
                double c2 = dR_du2[k] * srdc * uur * uu;
                double duur2 = e1n * e2m * srxc * uu * rijxu1[k], duu2 = e1n * e2m * srxc * uur * pr.u1[k];
                FC[k] = (-1.0) * (de2r + dRr + duur_r);
                TIc[k] = (-1.0) * (a1 + b1 + c1 + duur1 + duu1);
                TJc[k] = (-1.0) * (a2 + b2 + c2 + duur2 + duu2);
            }
            cross(pr.u1, TIc, T1c);
            cross(pr.u2, TJc, T2c);
            double potc = e1n * e2m * srxc * uur * uu;
            pot = pot + chirality * potc;
            for (int k = 0; k < 3; ++k) {
                FIJ[k] = FIJ[k] + chirality * FC[k];
                T1[k] = T1[k] + chirality * T1c[k];
                T2[k] = T2[k] + chirality * T2c[k];
            }
        }
        if (variant != 1 && (rij_mag >= switch_) && (rij_mag <= rlim)) {
            for (int k = 0; k < 3; ++k) {
                FIJ[k] = sx * FIJ[k] - sdx * pot * factor;
                T1[k] = sx * T1[k] - sdx * pot * factor;
                T2[k] = sx * T2[k] - sdx * pot * factor;
            }
            pot = sx * pot;
        }
    }
    *pe = pot;
    for (int k = 0; k < 3; ++k) { f[k] = FIJ[k]; ti[k] = T1[k]; tj[k] = T2[k]; }
}

void or_gb(const or_sys *s, int i, int j, int variant, double cutoff, int mu, int nu,
           double *pe, double f[3], double ti[3], double tj[3])
{
    or_gbx(s, i, j, variant, cutoff, 0.0, 0.0, mu, nu, pe, f, ti, tj);
}

/* ------------------------------------------------------------------ */
/* ECP (force.f90:1034-1551)                                           */
/* ------------------------------------------------------------------ */
double or_ecp_optim_dist(double l, const double ar[9], const double br[9], const double r[3])
{
    double gm[9], km[3];
    for (int k = 0; k < 9; ++k) gm[k] = (1 - l) * ar[k] + l * br[k];
    or_cholesky_solve(gm, r, km);
    return -l * (1 - l) * (r[0] * km[0] + r[1] * km[1] + r[2] * km[2]);
}

/* brent (force.f90:1046-1172), the scipy Brent.optimize control flow on
 * bracket (0, psi, 1) with atol 1e-3, rtol 1e-2 */
double or_ecp_brent(const double ar[9], const double br[9], const double r[3], int *niter)
{
    const double psi = 0.5 * (3.0 - sqrt(5.0));
    double a = 0.0, b = 1.0, deltax = 0.0, atol = 1e-3, rtol = 1e-2, rat = 0.0;
    int iter = 0;
    double c = (1 - psi) * a + psi * b;
    double fc = or_ecp_optim_dist(c, ar, br, r);
    double v = c, w = v, x = w;
    double fv = fc, fw = fv, ft = fw;
    while (iter < 10000) {
        double tol1 = rtol * fabs(x) + atol;
        double tol2 = 2.0 * tol1;
        double xmid = 0.5 * (a + b);
        if (fabs(x - xmid) < (tol2 - 0.5 * (b - a))) break;
        if (fabs(deltax) <= tol1) {
            deltax = (x >= xmid) ? a - x : b - x;
            rat = psi * deltax;
        } else {
            double tmp1 = (x - w) * (ft - fv);
            double tmp2 = (x - v) * (ft - fw);
            double p = (x - v) * tmp2 - (x - w) * tmp1;
            tmp2 = 2.0 * (tmp2 - tmp1);
            if (tmp2 > 0.0) p = -p;
            tmp2 = fabs(tmp2);
            double dx_temp = deltax;
            deltax = rat;
            if ((p > tmp2 * (a - x)) && (p < tmp2 * (b - x)) && (fabs(p) < fabs(0.5 * tmp2 * dx_temp))) {
                rat = p * 1.0 / tmp2;
                double u = x + rat;
                if ((u - a) < tol2 || (b - u) < tol2) rat = (xmid - x >= 0) ? tol1 : -tol1;
            } else {
                deltax = (x >= xmid) ? a - x : b - x;
                rat = psi * deltax;
            }
        }
        double u;
        if (fabs(rat) < tol1) u = (rat >= 0) ? x + tol1 : x - tol1;
        else u = x + rat;
        double fu = or_ecp_optim_dist(u, ar, br, r);
        if (fu > ft) {
            if (u < x) a = u; else b = u;
            if ((fu <= fw) || (w == x)) {
                v = w; w = u; fv = fw; fw = fu;
            } else if ((fu <= fv) || (v == x) || (v == w)) {
                v = u; fv = fu;
            }
        } else {
            if (u >= x) a = x; else b = x;
            v = w; w = x; x = u;
            fv = fw; fw = ft; ft = fu;
        }
        iter = iter + 1;
    }
    if (niter) *niter = iter;
    return x;
}

/* U = R(q)^T as built at force.f90:1244-1252 */
static void quat_U(const or_sys *s, int i, double U[9])
{
    double w = s->qw[i], x = s->qx[i], y = s->qy[i], z = s->qz[i];
    M(U, 1, 1) = w * w + x * x - y * y - z * z;
    M(U, 1, 2) = 2.0 * (x * y + w * z);
    M(U, 1, 3) = 2.0 * (x * z - w * y);
    M(U, 2, 1) = 2.0 * (x * y - w * z);
    M(U, 2, 2) = w * w - x * x + y * y - z * z;
    M(U, 2, 3) = 2.0 * (y * z + w * x);
    M(U, 3, 1) = 2.0 * (x * z + w * y);
    M(U, 3, 2) = 2.0 * (y * z - w * x);
    M(U, 3, 3) = w * w - x * x - y * y + z * z;
}

static double max3(double a, double b, double c) { return fmax(fmax(a, b), c); }

/* lambda_in < 0: the reference's Brent solve; otherwise the contact
   parameter is taken as given (used to evaluate the force expressions at a
   converged lambda*, where the envelope theorem makes F = -grad U exact). */
/* ecp_ (variant 0, force.f90:1202-1385), ecpr_ (1, :1387-1551), ecpw_ (2,
   welded: :1553-1742), ecpc_ (3, chiral: :1744-1984), ecpwc_ (4, welded and
   chiral: :1986-2234); lambda_in >= 0 replaces the Brent solve */
static void ecp_eval(const or_sys *s, int i, int j, int variant, double cutoff, double wf, double chirality,
                     double lambda_in, double *pe, double f[3], double ti[3], double tj[3],
                     double *lambda_out, int *niter_out)
{
    const int welded = (variant == 2 || variant == 4), chiral = (variant == 3 || variant == 4);
    const int meu = 2, neu = 1; /* force.f90:1038 */
    const double sigma_0 = 1.0;
    double pot = 0.0, FIJ[3] = {0, 0, 0}, T1[3] = {0, 0, 0}, T2[3] = {0, 0, 0};
    double RIJ[3];
    pair_rij(s, i, j, RIJ);
    double rij_sq = RIJ[0] * RIJ[0] + RIJ[1] * RIJ[1] + RIJ[2] * RIJ[2];
    double rij_mag = sqrt(rij_sq);
    double rhat[3] = {RIJ[0] / rij_mag, RIJ[1] / rij_mag, RIJ[2] / rij_mag};
    double U1[9], U2[9];
    quat_U(s, i, U1);
    quat_U(s, j, U2);
    double e10 = max3(s->eps_x[i], s->eps_y[i], s->eps_z[i]);
    double e20 = max3(s->eps_x[j], s->eps_y[j], s->eps_z[j]);
    double muinv = 1.0 / meu;
    double E1[9] = {0}, E2[9] = {0}, S1[9] = {0}, S2[9] = {0};
    M(E1, 1, 1) = pow(e10 / s->eps_x[i], muinv) / 4.0;
    M(E1, 2, 2) = pow(e10 / s->eps_y[i], muinv) / 4.0;
    M(E1, 3, 3) = pow(e10 / s->eps_z[i], muinv) / 4.0;
    M(E2, 1, 1) = pow(e20 / s->eps_x[j], muinv) / 4.0;
    M(E2, 2, 2) = pow(e20 / s->eps_y[j], muinv) / 4.0;
    M(E2, 3, 3) = pow(e20 / s->eps_z[j], muinv) / 4.0;
    M(S1, 1, 1) = s->sig_x[i] / 2.0; M(S1, 2, 2) = s->sig_y[i] / 2.0; M(S1, 3, 3) = s->sig_z[i] / 2.0;
    M(S2, 1, 1) = s->sig_x[j] / 2.0; M(S2, 2, 2) = s->sig_y[j] / 2.0; M(S2, 3, 3) = s->sig_z[j] / 2.0;
    double U1T[9], U2T[9], S1sq[9], S2sq[9];
    for (int a = 1; a <= 3; ++a)
        for (int b = 1; b <= 3; ++b) { M(U1T, a, b) = M(U1, b, a); M(U2T, a, b) = M(U2, b, a); }
    for (int k = 0; k < 9; ++k) { S1sq[k] = S1[k] * S1[k]; S2sq[k] = S2[k] * S2[k]; }
    double MM1[9], MM2[9], AR[9], BR[9], AE[9], BE[9];
    mmdot(U1T, S1, MM1);
    mmdot(U2T, S2, MM2);
    matmul_atba(U1, S1sq, AR);
    matmul_atba(U2, S2sq, BR);
    matmul_atba(U1, E1, AE);
    matmul_atba(U2, E2, BE);
    double st1 = (s->sig_x[i] * s->sig_y[i] + s->sig_z[i] * s->sig_z[i]) * sqrt(2.0 * s->sig_x[i] * s->sig_y[i]) / 8.0;
    double st2 = (s->sig_x[j] * s->sig_y[j] + s->sig_z[j] * s->sig_z[j]) * sqrt(2.0 * s->sig_x[j] * s->sig_y[j]) / 8.0;
    double sigma_t = sqrt(st1 * st2);
    double EAB[9], EM[9];
    for (int k = 0; k < 9; ++k) EAB[k] = AR[k] + BR[k];
    or_inverse(EAB, EM);
    double eps1 = sigma_t * sqrt(or_determinant(EM));
    int niter = 0;
    double lambda_R = lambda_in < 0.0 ? or_ecp_brent(AR, BR, RIJ, &niter) : lambda_in;
    double lambda_E = lambda_R;
    if (lambda_out) *lambda_out = lambda_R;
    if (niter_out) *niter_out = niter;
    double GE[9], GR[9], KE[3], KR[3];
    for (int k = 0; k < 9; ++k) GE[k] = (1 - lambda_E) * AE[k] + lambda_E * BE[k];
    or_cholesky_solve(GE, RIJ, KE);
    double eps2 = lambda_E * (1.0 - lambda_E) * (rhat[0] * KE[0] + rhat[1] * KE[1] + rhat[2] * KE[2]) / rij_mag;
    for (int k = 0; k < 9; ++k) GR[k] = (1 - lambda_R) * AR[k] + lambda_R * BR[k];
    or_cholesky_solve(GR, RIJ, KR);
    double phi = lambda_R * (1.0 - lambda_R) * (rhat[0] * KR[0] + rhat[1] * KR[1] + rhat[2] * KR[2]) / rij_mag;
    double sigma = 1.0 / sqrt(phi);
    double efac = 2.0 * lambda_E * (1.0 - lambda_E) / rij_sq;
    double rfac = 2.0 * lambda_R * (1.0 - lambda_R) / rij_sq;
    double sigma_0_inv = 1.0 / sigma_0;
    double phi_inv_3by2 = sigma * sigma * sigma;
    double switch_ = 0, factor = 0, sx = 0, sdx = 0, rlim;
    if (variant != 1) {
        rlim = welded ? (sigma + cutoff + wf - sigma_0) : (sigma + cutoff - sigma_0);
        switch_ = 0.9 * rlim;
        factor = 1.0 / (rlim - switch_);
        double xx = (rij_mag - switch_) * factor;
        sx = 1.0 - 6.0 * xx * xx * xx * xx * xx + 15.0 * xx * xx * xx * xx - 10.0 * xx * xx * xx;
        sdx = -30.0 * xx * xx * xx * xx + 60.0 * xx * xx * xx - 30.0 * xx * xx;
    } else {
        rlim = sigma;
    }
    int inside = (variant == 1) ? (rij_mag <= rlim) : (rij_mag < rlim);
    if (inside) {
        double sr;
        if (!welded) sr = sigma_0 / (rij_mag - sigma + sigma_0);
        else if (rij_mag < sigma) sr = sigma_0 / (rij_mag - sigma + sigma_0);
        else if (rij_mag >= sigma && rij_mag <= sigma + wf) sr = 1.0;
        else sr = sigma_0 / (rij_mag - sigma + sigma_0 - wf);
        double sr2 = sr * sr, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
        double srx = (sr12 - 2.0 * sr6);
        if (variant == 1) srx = srx + 1.0;
        double srd = 12.0 * (sr6 - sr12) * sr;
        double e1n = ipow(eps1, neu), e2m = ipow(eps2, meu), e2m1 = ipow(eps2, meu - 1);
        double dotKR = rhat[0] * KR[0] + rhat[1] * KR[1] + rhat[2] * KR[2];
        double dotKE = rhat[0] * KE[0] + rhat[1] * KE[1] + rhat[2] * KE[2];
        double dR_dr[3], de2_dr[3];
        for (int k = 0; k < 3; ++k) {
            if (chiral)
                dR_dr[k] = e1n * e2m * sigma_0_inv * (rhat[k] + 0.5 * rfac * (KR[k] - dotKR * rhat[k]) * phi_inv_3by2);
            else
                dR_dr[k] = sigma_0_inv * (rhat[k] + 0.5 * rfac * (KR[k] - dotKR * rhat[k]) * phi_inv_3by2);
            de2_dr[k] = e1n * (meu * e2m1) * efac * (KE[k] - dotKE * rhat[k]);
        }
        double de1_u1[3] = {0, 0, 0}, de1_u2[3] = {0, 0, 0};
        /* the chiral forms leave srx_fac / srd_fac out of the derivative terms */
        double pref1 = chiral ? e2m * (neu * e1n) : e2m * (neu * e1n) * srx;
        double d1x[3], d1y[3], d1z[3], tmp[3], col[3];
        for (int c = 0; c < 3; ++c) {
            for (int k = 0; k < 3; ++k) col[k] = MM1[k + 3 * c];
            mvdot(EM, col, tmp);
            double *dst = (c == 0) ? d1x : (c == 1) ? d1y : d1z;
            cross(col, tmp, dst);
            for (int k = 0; k < 3; ++k) dst[k] = pref1 * dst[k];
        }
        for (int k = 0; k < 3; ++k) de1_u1[k] = (d1x[k] + d1y[k]) + d1z[k];
        double KE_AE[3], KR_AR[3], ce[3], cr[3];
        vmdot(KE, AE, KE_AE);
        vmdot(KR, AR, KR_AR);
        cross(KE_AE, KE, ce);
        cross(KR_AR, KR, cr);
        double de2_u1[3], dR_u1[3];
        for (int k = 0; k < 3; ++k) {
            if (chiral) {
                de2_u1[k] = e1n * (meu * e2m1) * (1 - lambda_E) * efac * ce[k];
                dR_u1[k] = e1n * e2m * (0.5 * sigma_0_inv) * (1 - lambda_R) * rfac * cr[k] * phi_inv_3by2;
            } else {
                de2_u1[k] = e1n * (meu * e2m1) * srx * (1 - lambda_E) * efac * ce[k];
                dR_u1[k] = e1n * e2m * srd * (0.5 * sigma_0_inv) * (1 - lambda_R) * rfac * cr[k] * phi_inv_3by2;
            }
        }
        for (int c = 0; c < 3; ++c) {
            for (int k = 0; k < 3; ++k) col[k] = MM2[k + 3 * c];
            mvdot(EM, col, tmp);
            double *dst = (c == 0) ? d1x : (c == 1) ? d1y : d1z;
            cross(col, tmp, dst);
            for (int k = 0; k < 3; ++k) dst[k] = pref1 * dst[k];
        }
        for (int k = 0; k < 3; ++k) de1_u2[k] = (d1x[k] + d1y[k]) + d1z[k];
        double KE_BE[3], KR_BR[3];
        vmdot(KE, BE, KE_BE);
        vmdot(KR, BR, KR_BR);
        cross(KE_BE, KE, ce);
        cross(KR_BR, KR, cr);
        double de2_u2[3], dR_u2[3];
        for (int k = 0; k < 3; ++k) {
            if (chiral) {
                de2_u2[k] = e1n * (meu * e2m1) * lambda_E * efac * ce[k];
                dR_u2[k] = e1n * e2m * (0.5 * sigma_0_inv) * lambda_R * rfac * cr[k] * phi_inv_3by2;
            } else {
                de2_u2[k] = e1n * (meu * e2m1) * srx * lambda_E * efac * ce[k];
                dR_u2[k] = e1n * e2m * srd * (0.5 * sigma_0_inv) * lambda_R * rfac * cr[k] * phi_inv_3by2;
            }
        }
        pot = e1n * e2m * srx;
        for (int k = 0; k < 3; ++k) {
            if (chiral) {
                FIJ[k] = (-1.0) * (srx * de2_dr[k] + srd * dR_dr[k]);
                T1[k] = (srx * (de1_u1[k] + de2_u1[k]) + srd * dR_u1[k]);
                T2[k] = (srx * (de1_u2[k] + de2_u2[k]) + srd * dR_u2[k]);
            } else {
                FIJ[k] = (-1.0) * (e1n * e2m * srd * dR_dr[k] + srx * de2_dr[k]);
                T1[k] = (de1_u1[k] + de2_u1[k]) + dR_u1[k];
                T2[k] = (de1_u2[k] + de2_u2[k]) + dR_u2[k];
            }
        }
        if (chiral) {
            /* force.f90:1916-1968 (ecpc_), :2170-2222 (ecpwc_) */
            double srxc, srdc;
            if (welded) {
                srxc = (sr12 - (12.0 / 7.0) * sr6 * sr);
                srdc = 12.0 * (sr6 * sr2 - sr12 * sr);
            } else {
                srxc = sr6 * sr;
                srdc = -7.0 * sr6 * sr2;
            }
            double u1z[3] = {M(U1, 3, 1), M(U1, 3, 2), M(U1, 3, 3)};
            double u2z[3] = {M(U2, 3, 1), M(U2, 3, 2), M(U2, 3, 3)};
            double u1xu2[3], u2xrij[3], rijxu1[3];
            cross(u1z, u2z, u1xu2);
            cross(u2z, rhat, u2xrij);
            cross(rhat, u1z, rijxu1);
            double uur = u1xu2[0] * rhat[0] + u1xu2[1] * rhat[1] + u1xu2[2] * rhat[2];
            double uu = u1z[0] * u2z[0] + u1z[1] * u2z[1] + u1z[2] * u2z[2];
            double term1[3], term2[3], tu1[3], tu2[3], FC[3], T1c[3], T2c[3];
            for (int k = 0; k < 3; ++k) {
                double duur1 = e1n * e2m * srxc * uu * u2xrij[k], duu1 = e1n * e2m * srxc * uur * u2z[k];
                double duur2 = e1n * e2m * srxc * uu * rijxu1[k], duu2 = e1n * e2m * srxc * uur * u1z[k];
                term1[k] = (duur1 + duu1);
                term2[k] = (duur2 + duu2);
            }
            cross(u1z, term1, tu1);
            cross(u2z, term2, tu2);
            for (int k = 0; k < 3; ++k) {
                double dRr = dR_dr[k] * srdc * uur * uu;
                double de2r = de2_dr[k] * srxc * uur * uu;
                double duur_r = e1n * e2m * srxc * uu * (u1xu2[k] - uur * rhat[k]) / rij_mag;
                double a1 = de1_u1[k] * srxc * uur * uu, b1 = de2_u1[k] * srxc * uur * uu;
                double c1 = dR_u1[k] * srdc * uur * uu;
                double a2 = de1_u2[k] * srxc * uur * uu, b2 = de2_u2[k] * srxc * uur * uu;
                double c2 = dR_u2[k] * srdc * uur * uu;
                FC[k] = (-1.0) * (de2r + dRr + duur_r);
                T1c[k] = (a1 + b1 + c1 - tu1[k]);
                T2c[k] = (a2 + b2 + c2 - tu2[k]);
            }
            double potc = e1n * e2m * srxc * uur * uu;
            pot = pot + chirality * potc;
            for (int k = 0; k < 3; ++k) {
                FIJ[k] = FIJ[k] + chirality * FC[k];
                T1[k] = T1[k] + chirality * T1c[k];
                T2[k] = T2[k] + chirality * T2c[k];
            }
        }
        if (variant != 1 && (rij_mag >= switch_) && (rij_mag <= rlim)) {
            for (int k = 0; k < 3; ++k) {
                FIJ[k] = sx * FIJ[k] - sdx * pot * factor;
                T1[k] = sx * T1[k] - sdx * pot * factor;
                T2[k] = sx * T2[k] - sdx * pot * factor;
            }
            pot = sx * pot;
        }
    }
    *pe = pot;
    for (int k = 0; k < 3; ++k) { f[k] = FIJ[k]; ti[k] = T1[k]; tj[k] = T2[k]; }
}

void or_ecp(const or_sys *s, int i, int j, int variant, double cutoff,
            double *pe, double f[3], double ti[3], double tj[3],
            double *lambda_out, int *niter_out)
{
    ecp_eval(s, i, j, variant, cutoff, 0.0, 0.0, -1.0, pe, f, ti, tj, lambda_out, niter_out);
}

void or_ecpx(const or_sys *s, int i, int j, int variant, double cutoff, double wf, double chirality,
             double *pe, double f[3], double ti[3], double tj[3], double *lambda_out, int *niter_out)
{
    ecp_eval(s, i, j, variant, cutoff, wf, chirality, -1.0, pe, f, ti, tj, lambda_out, niter_out);
}

void or_ecp_at_lambda(const or_sys *s, int i, int j, int variant, double cutoff, double lambda,
                      double *pe, double f[3], double ti[3], double tj[3])
{
    ecp_eval(s, i, j, variant, cutoff, 0.0, 0.0, lambda, pe, f, ti, tj, NULL, NULL);
}

void or_ecpx_at_lambda(const or_sys *s, int i, int j, int variant, double cutoff, double wf, double chirality,
                       double lambda, double *pe, double f[3], double ti[3], double tj[3])
{
    ecp_eval(s, i, j, variant, cutoff, wf, chirality, lambda, pe, f, ti, tj, NULL, NULL);
}

/* the pair's shape matrices AR, BR (force.f90:1254-1259) and min-image
   separation, so a caller can maximise the contact function itself */
void or_ecp_shape(const or_sys *s, int i, int j, double ar[9], double br[9], double rij[3])
{
    double U1[9], U2[9], S1sq[9] = {0}, S2sq[9] = {0};
    pair_rij(s, i, j, rij);
    quat_U(s, i, U1);
    quat_U(s, j, U2);
    M(S1sq, 1, 1) = (s->sig_x[i] / 2.0) * (s->sig_x[i] / 2.0);
    M(S1sq, 2, 2) = (s->sig_y[i] / 2.0) * (s->sig_y[i] / 2.0);
    M(S1sq, 3, 3) = (s->sig_z[i] / 2.0) * (s->sig_z[i] / 2.0);
    M(S2sq, 1, 1) = (s->sig_x[j] / 2.0) * (s->sig_x[j] / 2.0);
    M(S2sq, 2, 2) = (s->sig_y[j] / 2.0) * (s->sig_y[j] / 2.0);
    M(S2sq, 3, 3) = (s->sig_z[j] / 2.0) * (s->sig_z[j] / 2.0);
    matmul_atba(U1, S1sq, ar);
    matmul_atba(U2, S2sq, br);
}

/* ---- converged contact parameter and network features (no reference
   counterpart: the converged and learned ECP solver modes) ---- */

/* F(l) = l(1-l) r^T G^-1 r with G = (1-l) A + l B, and dF/dl, d2F/dl2 */
static void ecp_F_derivs(double l, const double ar[9], const double br[9], const double r[3],
                         double *F, double *d1, double *d2)
{
    double G[9], D[9], k[3], u[3], w[3];
    for (int q = 0; q < 9; ++q) { G[q] = (1 - l) * ar[q] + l * br[q]; D[q] = br[q] - ar[q]; }
    or_cholesky_solve(G, r, k);
    for (int a = 0; a < 3; ++a) u[a] = D[a] * k[0] + D[a + 3] * k[1] + D[a + 6] * k[2];
    or_cholesky_solve(G, u, w);
    double rk = r[0] * k[0] + r[1] * k[1] + r[2] * k[2];
    double kdk = k[0] * u[0] + k[1] * u[1] + k[2] * u[2];
    double udw = u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
    double ll = l * (1 - l);
    *F = ll * rk;
    *d1 = (1 - 2 * l) * rk - ll * kdk;
    *d2 = -2 * rk - 2 * (1 - 2 * l) * kdk + 2 * ll * udw;
}

/* maximiser of F on (0,1) (F is concave there) by Newton on F' = 0 from l0:
   a step below tol ends the solve; a step that leaves the bracket [lo, hi]
   kept by the sign of F' is replaced by bisection */
double or_ecp_lambda_converged(const double ar[9], const double br[9], const double r[3], double l0, double tol,
                               int *iters)
{
    double lo = 0.0, hi = 1.0, l = fmin(fmax(l0, 1e-6), 1.0 - 1e-6);
    int it = 0;
    for (; it < 100; ++it) {
        double F, d1, d2;
        ecp_F_derivs(l, ar, br, r, &F, &d1, &d2);
        if (d1 > 0) lo = l; else hi = l;
        if (d2 < 0) {
            double step = -d1 / d2;
            if (fabs(step) < tol) { l = l + step; ++it; break; }
            double ln = l + step;
            l = (ln > lo && ln < hi) ? ln : 0.5 * (lo + hi);
        } else {
            l = 0.5 * (lo + hi);
        }
        if (hi - lo < tol) { ++it; break; }
    }
    if (iters) *iters = it;
    return l;
}

/* contact distance sigma = 1/sqrt(l(1-l) rhat^T G(l)^-1 rhat) */
double or_ecp_contact_distance(const double ar[9], const double br[9], const double r[3], double l)
{
    double G[9], k[3];
    double m = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    double rh[3] = {r[0] / m, r[1] / m, r[2] / m};
    for (int q = 0; q < 9; ++q) G[q] = (1 - l) * ar[q] + l * br[q];
    or_cholesky_solve(G, rh, k);
    return 1.0 / sqrt(l * (1 - l) * (rh[0] * k[0] + rh[1] * k[1] + rh[2] * k[2]));
}

/* network inputs of a pair: the unit separation rhat = RIJ/|RIJ| in particle
   i's body frame (U_i rhat, U = R(q)^T) and the relative rotation
   conj(q_i) q_j with its sign chosen so that w >= 0 */
void or_ecp_features(const or_sys *s, int i, int j, double x[7])
{
    double RIJ[3], U[9];
    pair_rij(s, i, j, RIJ);
    double m = sqrt(RIJ[0] * RIJ[0] + RIJ[1] * RIJ[1] + RIJ[2] * RIJ[2]);
    double rh[3] = {RIJ[0] / m, RIJ[1] / m, RIJ[2] / m};
    quat_U(s, i, U);
    for (int a = 1; a <= 3; ++a) x[a - 1] = M(U, a, 1) * rh[0] + M(U, a, 2) * rh[1] + M(U, a, 3) * rh[2];
    double aw = s->qw[i], ax = -s->qx[i], ay = -s->qy[i], az = -s->qz[i];
    double bw = s->qw[j], bx = s->qx[j], by = s->qy[j], bz = s->qz[j];
    double w = aw * bw - ax * bx - ay * by - az * bz;
    double qx = aw * bx + ax * bw + ay * bz - az * by;
    double qy = aw * by - ax * bz + ay * bw + az * bx;
    double qz = aw * bz + ax * by - ay * bx + az * bw;
    double sg = (w < 0) ? -1.0 : 1.0;
    x[3] = sg * w; x[4] = sg * qx; x[5] = sg * qy; x[6] = sg * qz;
}

/* ------------------------------------------------------------------ */
/* bonded (force.f90:2244-2462)                                        */
/* ------------------------------------------------------------------ */
static inline void bvec(const or_sys *s, int a, int b, double v[3])
{
    /* V = R(b) - R(a) with min image */
    v[0] = s->rx[b] - s->rx[a];
    v[1] = s->ry[b] - s->ry[a];
    v[2] = s->rz[b] - s->rz[a];
    if (s->periodic) {
        v[0] = mimg(v[0], s->box[0]);
        v[1] = mimg(v[1], s->box[1]);
        v[2] = mimg(v[2], s->box[2]);
    }
}

static void bend_common(const or_sys *s, int i, int j, int k, double strength, int mode,
                        double coeff, double angle, double *pe,
                        double fi[3], double fj[3], double fk[3])
{
    double VI[3], VJ[3];
    bvec(s, i, j, VI);
    bvec(s, j, k, VJ);
    double CC11 = VI[0] * VI[0] + VI[1] * VI[1] + VI[2] * VI[2];
    double CC12 = VI[0] * VJ[0] + VI[1] * VJ[1] + VI[2] * VJ[2];
    double CC22 = VJ[0] * VJ[0] + VJ[1] * VJ[1] + VJ[2] * VJ[2];
    double prefac = 1.0 / sqrt(CC11 * CC22);
    double fac = CC12, fac1 = fac / CC22, fac2 = fac / CC11;
    if (mode == 0) { /* bend_ : k (1 - cos) */
        *pe = coeff * (1.0 - prefac * fac);
        strength = coeff;
    } else if (mode == 2) { /* angle_bend_ (force.f90:2464-2506): k/2 (theta - theta0)^2 */
        double theta0 = PI_REF - (PI_REF / 180.0) * angle;
        double theta = acos(prefac * fac);
        strength = coeff * (theta - theta0) / sin(theta);
        *pe = (coeff / 2.0) * (theta - theta0) * (theta - theta0);
    } else {         /* cosine_bend_ */
        double theta0 = PI_REF - (PI_REF / 180.0) * angle;
        double cos_theta0 = cos(theta0);
        double cos_theta = prefac * fac;
        strength = coeff * (cos_theta - cos_theta0);
        *pe = (coeff / 2.0) * (cos_theta - cos_theta0) * (cos_theta - cos_theta0);
    }
    for (int q = 0; q < 3; ++q) {
        fk[q] = -strength * prefac * (fac1 * VJ[q] - VI[q]);
        fj[q] = strength * prefac * (fac1 * VJ[q] - fac2 * VI[q] + VJ[q] - VI[q]);
        fi[q] = strength * prefac * (fac2 * VI[q] - VJ[q]);
    }
}

void or_bend(const or_sys *s, int i, int j, int k, double coeff,
             double *pe, double fi[3], double fj[3], double fk[3])
{
    bend_common(s, i, j, k, 0.0, 0, coeff, 0.0, pe, fi, fj, fk);
}

void or_cosine_bend(const or_sys *s, int i, int j, int k, double coeff, double angle,
                    double *pe, double fi[3], double fj[3], double fk[3])
{
    bend_common(s, i, j, k, 0.0, 1, coeff, angle, pe, fi, fj, fk);
}

static void dih_common(const or_sys *s, int i, int j, int k, int l, int mode,
                       double coeff, double angle, double *pe,
                       double fi[3], double fj[3], double fk[3], double fl[3])
{
    double VI[3], VJ[3], VK[3];
    bvec(s, i, j, VI);
    bvec(s, j, k, VJ);
    bvec(s, k, l, VK);
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
    double strength;
    if (mode == 0) { /* dih_ : k (1 + prefac*fac) */
        *pe = coeff * (1.0 + prefac * fac);
        strength = coeff;
    } else if (mode == 2) { /* angle_dih_ (force.f90:2508-2579): k (1 - cos(phi - phi0)) */
        double vixvj[3];
        cross(VI, VJ, vixvj);
        double fac_sin = vixvj[0] * VK[0] + vixvj[1] * VK[1] + vixvj[2] * VK[2];
        double phi0 = PI_REF - (PI_REF / 180.0) * angle;
        double cos_phi0 = cos(phi0), sin_phi0 = sin(phi0);
        double cos_phi = -prefac * fac;
        double sin_phi = prefac * fac_sin * sqrt(VJ[0] * VJ[0] + VJ[1] * VJ[1] + VJ[2] * VJ[2]);
        strength = coeff * (cos_phi0 - sin_phi0 * (cos_phi / sin_phi));
        *pe = coeff * (1.0 - cos_phi * cos_phi0 - sin_phi * sin_phi0);
    } else {         /* cosine_dih_ */
        double vixvj[3];
        cross(VI, VJ, vixvj);
        double fac_sin = vixvj[0] * VK[0] + vixvj[1] * VK[1] + vixvj[2] * VK[2];
        double phi0 = PI_REF - (PI_REF / 180.0) * angle;
        double cos_phi0 = cos(phi0);
        double cos_phi = -prefac * fac;
        double sin_phi = prefac * fac_sin * sqrt(VJ[0] * VJ[0] + VJ[1] * VJ[1] + VJ[2] * VJ[2]);
        double sgn_phi = atan2(sin_phi, cos_phi);
        strength = coeff * (cos(sgn_phi) - cos_phi0);
        *pe = (coeff / 2.0) * (cos(sgn_phi) - cos_phi0) * (cos(sgn_phi) - cos_phi0);
    }
    for (int q = 0; q < 3; ++q) {
        fl[q] = -strength * prefac * (CC12 * VJ[q] - CC22 * VI[q] - fac1 * (CC22 * VK[q] - CC23 * VJ[q]));
        fk[q] = -strength * prefac * (CC12 * VK[q] - CC12 * VJ[q] + CC23 * VI[q] + CC22 * VI[q] - 2.0 * CC13 * VJ[q]
                                      - fac2 * (CC11 * VJ[q] - CC12 * VI[q]) - fac1 * (CC33 * VJ[q] - CC22 * VK[q] - CC23 * VK[q] + CC23 * VJ[q]));
        fj[q] = -strength * prefac * (-CC12 * VK[q] + CC23 * VJ[q] - CC23 * VI[q] - CC22 * VK[q] + 2.0 * CC13 * VJ[q]
                                      - fac2 * (CC22 * VI[q] - CC11 * VJ[q] - CC12 * VJ[q] + CC12 * VI[q]) - fac1 * (-CC33 * VJ[q] + CC23 * VK[q]));
        fi[q] = -strength * prefac * (-CC23 * VJ[q] + CC22 * VK[q] - fac2 * (-CC22 * VI[q] + CC12 * VJ[q]));
    }
}

void or_angle_bend(const or_sys *s, int i, int j, int k, double coeff, double angle,
                   double *pe, double fi[3], double fj[3], double fk[3])
{
    bend_common(s, i, j, k, 0.0, 2, coeff, angle, pe, fi, fj, fk);
}

void or_dih(const or_sys *s, int i, int j, int k, int l, double coeff,
            double *pe, double fi[3], double fj[3], double fk[3], double fl[3])
{
    dih_common(s, i, j, k, l, 0, coeff, 0.0, pe, fi, fj, fk, fl);
}

void or_cosine_dih(const or_sys *s, int i, int j, int k, int l, double coeff, double angle,
                   double *pe, double fi[3], double fj[3], double fk[3], double fl[3])
{
    dih_common(s, i, j, k, l, 1, coeff, angle, pe, fi, fj, fk, fl);
}

void or_angle_dih(const or_sys *s, int i, int j, int k, int l, double coeff, double angle,
                  double *pe, double fi[3], double fj[3], double fk[3], double fl[3])
{
    dih_common(s, i, j, k, l, 2, coeff, angle, pe, fi, fj, fk, fl);
}

/* ------------------------------------------------------------------ */
/* integrators (integrators.f90)                                       */
/* ------------------------------------------------------------------ */
static inline void wrap_pos(or_sys *s, int i)
{
    if (s->periodic) {
        s->rx[i] = s->rx[i] - anint(s->rx[i] / s->box[0]) * s->box[0];
        s->ry[i] = s->ry[i] - anint(s->ry[i] / s->box[1]) * s->box[1];
        s->rz[i] = s->rz[i] - anint(s->rz[i] / s->box[2]) * s->box[2];
    }
}

static inline void half_kick(or_sys *s, int i, double dt)
{
    s->vx[i] = s->vx[i] + 0.5 * dt * (s->fx[i] * s->inv_mass[i]);
    s->vy[i] = s->vy[i] + 0.5 * dt * (s->fy[i] * s->inv_mass[i]);
    s->vz[i] = s->vz[i] + 0.5 * dt * (s->fz[i] * s->inv_mass[i]);
}

static inline void save_old(or_sys *s, int i)
{
    s->rx_old[i] = s->rx[i]; s->ry_old[i] = s->ry[i]; s->rz_old[i] = s->rz[i];
}

static void vv_initial(or_sys *s, double dt, int save)
{
    for (int i = 0; i < s->n; ++i) {
        half_kick(s, i, dt);
        if (save) save_old(s, i);
        s->rx[i] = s->rx[i] + dt * s->vx[i];
        s->ry[i] = s->ry[i] + dt * s->vy[i];
        s->rz[i] = s->rz[i] + dt * s->vz[i];
        wrap_pos(s, i);
    }
}

static void vv_final(or_sys *s, double dt)
{
    s->kin_eng = 0.0;
    for (int i = 0; i < s->n; ++i) {
        half_kick(s, i, dt);
        s->kin_eng = s->kin_eng + 0.5 * s->mass[i] * (s->vx[i] * s->vx[i] + s->vy[i] * s->vy[i] + s->vz[i] * s->vz[i]);
    }
}

void or_vv_initial_step(or_sys *s, double dt) { vv_initial(s, dt, 0); }
void or_vv_final_step(or_sys *s, double dt) { vv_final(s, dt); }
void or_vvc_initial_step(or_sys *s, double dt) { vv_initial(s, dt, 1); or_apply_constraints_a(s, dt); }
void or_vvc_final_step(or_sys *s, double dt) { vv_final(s, dt); or_apply_constraints_b(s, dt); }

/* RM = R(q)^T rows as at integrators.f90:144-152 */
static inline void rm_of(double w, double x, double y, double z, double RM[9])
{
    M(RM, 1, 1) = w * w + x * x - y * y - z * z;
    M(RM, 1, 2) = 2.0 * (x * y + w * z);
    M(RM, 1, 3) = 2.0 * (x * z - w * y);
    M(RM, 2, 1) = 2.0 * (x * y - w * z);
    M(RM, 2, 2) = w * w - x * x + y * y - z * z;
    M(RM, 2, 3) = 2.0 * (y * z + w * x);
    M(RM, 3, 1) = 2.0 * (x * z + w * y);
    M(RM, 3, 2) = 2.0 * (y * z - w * x);
    M(RM, 3, 3) = w * w - x * x - y * y + z * z;
}

/* qq_initial_step (integrators.f90:123-183) */
void or_qq_initial_step(or_sys *s, double dt)
{
    double RM[9];
    for (int i = 0; i < s->n; ++i) {
        double qwn = s->qw[i], qxn = s->qx[i], qyn = s->qy[i], qzn = s->qz[i];
        double qdw = 0, qdx = 0, qdy = 0, qdz = 0;
        s->lx[i] = s->lx[i] + 0.5 * dt * s->tx[i];
        s->ly[i] = s->ly[i] + 0.5 * dt * s->ty[i];
        s->lz[i] = s->lz[i] + 0.5 * dt * s->tz[i];
        for (int pass = 0; pass < 2; ++pass) {
            double qwo = qwn, qxo = qxn, qyo = qyn, qzo = qzn;
            rm_of(qwo, qxo, qyo, qzo, RM);
            double wx = (M(RM, 1, 1) * s->lx[i] + M(RM, 1, 2) * s->ly[i] + M(RM, 1, 3) * s->lz[i]) * s->ixx_inv[i];
            double wy = (M(RM, 2, 1) * s->lx[i] + M(RM, 2, 2) * s->ly[i] + M(RM, 2, 3) * s->lz[i]) * s->iyy_inv[i];
            double wz = (M(RM, 3, 1) * s->lx[i] + M(RM, 3, 2) * s->ly[i] + M(RM, 3, 3) * s->lz[i]) * s->izz_inv[i];
            qdw = 0.5 * (-qxo * wx - qyo * wy - qzo * wz);
            qdx = 0.5 * (qwo * wx + qyo * wz - qzo * wy);
            qdy = 0.5 * (qwo * wy - qxo * wz + qzo * wx);
            qdz = 0.5 * (qwo * wz + qxo * wy - qyo * wx);
            qwn = s->qw[i] + 0.5 * dt * qdw;
            qxn = s->qx[i] + 0.5 * dt * qdx;
            qyn = s->qy[i] + 0.5 * dt * qdy;
            qzn = s->qz[i] + 0.5 * dt * qdz;
        }
        s->qw[i] = s->qw[i] + dt * qdw;
        s->qx[i] = s->qx[i] + dt * qdx;
        s->qy[i] = s->qy[i] + dt * qdy;
        s->qz[i] = s->qz[i] + dt * qdz;
        double qm = sqrt(s->qw[i] * s->qw[i] + s->qx[i] * s->qx[i] + s->qy[i] * s->qy[i] + s->qz[i] * s->qz[i]);
        s->qw[i] = s->qw[i] / qm; s->qx[i] = s->qx[i] / qm;
        s->qy[i] = s->qy[i] / qm; s->qz[i] = s->qz[i] / qm;
    }
}

static inline double kin_rot_of(const or_sys *s, int i, double RM[9])
{
    rm_of(s->qw[i], s->qx[i], s->qy[i], s->qz[i], RM);
    double lbx = (M(RM, 1, 1) * s->lx[i] + M(RM, 1, 2) * s->ly[i] + M(RM, 1, 3) * s->lz[i]);
    double lby = (M(RM, 2, 1) * s->lx[i] + M(RM, 2, 2) * s->ly[i] + M(RM, 2, 3) * s->lz[i]);
    double lbz = (M(RM, 3, 1) * s->lx[i] + M(RM, 3, 2) * s->ly[i] + M(RM, 3, 3) * s->lz[i]);
    return 0.5 * ((lbx * lbx) * s->ixx_inv[i] + (lby * lby) * s->iyy_inv[i] + (lbz * lbz) * s->izz_inv[i]);
}

/* qq_final_step (integrators.f90:185-218) */
void or_qq_final_step(or_sys *s, double dt)
{
    double RM[9];
    s->kin_rot = 0.0;
    for (int i = 0; i < s->n; ++i) {
        s->lx[i] = s->lx[i] + 0.5 * dt * s->tx[i];
        s->ly[i] = s->ly[i] + 0.5 * dt * s->ty[i];
        s->lz[i] = s->lz[i] + 0.5 * dt * s->tz[i];
        s->kin_rot = s->kin_rot + kin_rot_of(s, i, RM);
    }
}

/* ---- Nose-Hoover chain (integrators.f90:222-492) ---- */
static double polyval(double x, const double *c, int upper)
{
    double f = c[upper];
    for (int i = upper - 1; i >= 0; --i) f = f * x + c[i];
    return f;
}

double or_exprel(double x)
{
    const double g[5] = {1.0 / 1, 1.0 / 6, 1.0 / 120, 1.0 / 5040, 1.0 / 362880};
    const double tol = 0.01;
    if (fabs(x) > tol) return (exp(x) - 1.0) / x;
    double h = x / 2;
    return exp(x / 2) * polyval(h * h, g, 4);
}

static void u4_propagator(or_sys *s, double t, int j_start, int j_stop, int j_stride)
{
    double n_dof = (double)(3 * (s->n / s->nrigid) - 3);
    for (int j = j_start; (j_stride > 0) ? (j <= j_stop) : (j >= j_stop); j += j_stride) {
        double gj;
        if (j == 1) {
            double sum = 0.0;
            for (int i = 0; i < s->n; ++i)
                sum += s->mass[i] * (s->vx[i] * s->vx[i] + s->vy[i] * s->vy[i] + s->vz[i] * s->vz[i]);
            gj = sum - n_dof * s->target_temp;
        } else {
            gj = (s->p_eta[j - 2] * s->p_eta[j - 2] / s->hb[j - 2]) - s->target_temp;
        }
        if (j == 3) {
            s->p_eta[j - 1] = s->p_eta[j - 1] + t * gj;
        } else {
            double x = t * s->p_eta[j] / s->hb[j];
            double c = or_exprel(-x);
            s->p_eta[j - 1] = s->p_eta[j - 1] * exp(-x) + t * gj * c;
        }
    }
}

static void nh_set_hb(or_sys *s, const double *coupling)
{
    double n_dof = (double)(3 * (s->n / s->nrigid) - 3);
    const double tau = 2.0;
    if (coupling) {
        s->hb[0] = s->hb[1] = s->hb[2] = *coupling;
    } else {
        s->hb[0] = s->hb[1] = s->hb[2] = s->target_temp * (tau * tau);
        s->hb[0] = n_dof * s->target_temp * (tau * tau);
    }
}

static inline void nh_scale(or_sys *s, int i, double dt)
{
    s->vx[i] = s->vx[i] * exp(-0.5 * dt * s->p_eta[0] / s->hb[0]);
    s->vy[i] = s->vy[i] * exp(-0.5 * dt * s->p_eta[0] / s->hb[0]);
    s->vz[i] = s->vz[i] * exp(-0.5 * dt * s->p_eta[0] / s->hb[0]);
    for (int k = 0; k < 3; ++k) s->eta[k] = s->eta[k] + 0.5 * dt * s->p_eta[k] / s->hb[k];
}

static void nht_initial(or_sys *s, double dt, const double *coupling, int save)
{
    nh_set_hb(s, coupling);
    for (int i = 0; i < s->n; ++i) {
        u4_propagator(s, dt / 4.0, 3, 1, -1);
        nh_scale(s, i, dt);
        u4_propagator(s, dt / 4.0, 1, 3, 1);
        half_kick(s, i, dt);
        if (save) save_old(s, i);
        s->rx[i] = s->rx[i] + dt * s->vx[i];
        s->ry[i] = s->ry[i] + dt * s->vy[i];
        s->rz[i] = s->rz[i] + dt * s->vz[i];
        wrap_pos(s, i);
    }
}

static void nht_final(or_sys *s, double dt, const double *coupling)
{
    nh_set_hb(s, coupling);
    s->kin_eng = 0.0;
    for (int i = 0; i < s->n; ++i) {
        half_kick(s, i, dt);
        u4_propagator(s, dt / 4.0, 3, 1, -1);
        nh_scale(s, i, dt);
        u4_propagator(s, dt / 4.0, 1, 3, 1);
        s->kin_eng = s->kin_eng + 0.5 * s->mass[i] * (s->vx[i] * s->vx[i] + s->vy[i] * s->vy[i] + s->vz[i] * s->vz[i]);
    }
}

void or_nht_initial_step(or_sys *s, double dt, const double *c) { nht_initial(s, dt, c, 0); }
void or_nht_final_step(or_sys *s, double dt, const double *c) { nht_final(s, dt, c); }
void or_nhtc_initial_step(or_sys *s, double dt, const double *c) { nht_initial(s, dt, c, 1); or_apply_constraints_a(s, dt); }
void or_nhtc_final_step(or_sys *s, double dt, const double *c) { nht_final(s, dt, c); or_apply_constraints_b(s, dt); }

/* ---- Langevin (integrators.f90:494-857) ---- */
static void lgv_initial(or_sys *s, double dt, int step, double damp, int save)
{
    double kk1 = 1.0 - (damp * dt) / 2.0;
    double kk2 = 1.0 / (1.0 + (damp * dt) / 2.0);
    int temp_int = (int)(s->target_temp * 1000.0);
    double T = s->target_temp;
    double RM[9];
    for (int i = 0; i < s->n; ++i) {
        double n1x, n1y, n1z, n2x, n2y, n2z;
        half_kick(s, i, dt);
        if (save) save_old(s, i);
        s->rx[i] = s->rx[i] + 0.5 * dt * s->vx[i];
        s->ry[i] = s->ry[i] + 0.5 * dt * s->vy[i];
        s->rz[i] = s->rz[i] + 0.5 * dt * s->vz[i];
        or_philox_normal_3seq(i + 1, step, temp_int, 0, &n1x, &n1y, &n1z);
        or_philox_normal_3seq(i + 1, step, temp_int, 1, &n2x, &n2y, &n2z);
        s->vx[i] = exp(-damp * dt) * s->vx[i] + sqrt(T * s->inv_mass[i] * (1 - exp(-2.0 * damp * dt))) * n1x;
        s->vy[i] = exp(-damp * dt) * s->vy[i] + sqrt(T * s->inv_mass[i] * (1 - exp(-2.0 * damp * dt))) * n1y;
        s->vz[i] = exp(-damp * dt) * s->vz[i] + sqrt(T * s->inv_mass[i] * (1 - exp(-2.0 * damp * dt))) * n1z;
        s->rx[i] = s->rx[i] + 0.5 * dt * s->vx[i];
        s->ry[i] = s->ry[i] + 0.5 * dt * s->vy[i];
        s->rz[i] = s->rz[i] + 0.5 * dt * s->vz[i];
        wrap_pos(s, i);
        double qwn = s->qw[i], qxn = s->qx[i], qyn = s->qy[i], qzn = s->qz[i];
        s->lx[i] = s->lx[i] + 0.5 * dt * s->tx[i];
        s->ly[i] = s->ly[i] + 0.5 * dt * s->ty[i];
        s->lz[i] = s->lz[i] + 0.5 * dt * s->tz[i];
        double lbx = 0, lby = 0, lbz = 0, qdw = 0, qdx = 0, qdy = 0, qdz = 0;
        for (int pass = 0; pass < 2; ++pass) {
            double qwo = qwn, qxo = qxn, qyo = qyn, qzo = qzn;
            rm_of(qwo, qxo, qyo, qzo, RM);
            lbx = (M(RM, 1, 1) * s->lx[i] + M(RM, 1, 2) * s->ly[i] + M(RM, 1, 3) * s->lz[i]);
            lby = (M(RM, 2, 1) * s->lx[i] + M(RM, 2, 2) * s->ly[i] + M(RM, 2, 3) * s->lz[i]);
            lbz = (M(RM, 3, 1) * s->lx[i] + M(RM, 3, 2) * s->ly[i] + M(RM, 3, 3) * s->lz[i]);
            double wx = lbx * s->ixx_inv[i], wy = lby * s->iyy_inv[i], wz = lbz * s->izz_inv[i];
            qdw = 0.5 * (-qxo * wx - qyo * wy - qzo * wz);
            qdx = 0.5 * (qwo * wx + qyo * wz - qzo * wy);
            qdy = 0.5 * (qwo * wy - qxo * wz + qzo * wx);
            qdz = 0.5 * (qwo * wz + qxo * wy - qyo * wx);
            qwn = s->qw[i] + 0.5 * dt * qdw;
            qxn = s->qx[i] + 0.5 * dt * qdx;
            qyn = s->qy[i] + 0.5 * dt * qdy;
            qzn = s->qz[i] + 0.5 * dt * qdz;
        }
        lbx = kk2 * kk1 * lbx + kk2 * sqrt(2.0 * damp * T * dt * s->ixx[i]) * n2x;
        lby = kk2 * kk1 * lby + kk2 * sqrt(2.0 * damp * T * dt * s->iyy[i]) * n2y;
        lbz = kk2 * kk1 * lbz + kk2 * sqrt(2.0 * damp * T * dt * s->izz[i]) * n2z;
        s->lx[i] = (M(RM, 1, 1) * lbx + M(RM, 2, 1) * lby + M(RM, 3, 1) * lbz);
        s->ly[i] = (M(RM, 1, 2) * lbx + M(RM, 2, 2) * lby + M(RM, 3, 2) * lbz);
        s->lz[i] = (M(RM, 1, 3) * lbx + M(RM, 2, 3) * lby + M(RM, 3, 3) * lbz);
        s->qw[i] = s->qw[i] + dt * qdw;
        s->qx[i] = s->qx[i] + dt * qdx;
        s->qy[i] = s->qy[i] + dt * qdy;
        s->qz[i] = s->qz[i] + dt * qdz;
        double qm = sqrt(s->qw[i] * s->qw[i] + s->qx[i] * s->qx[i] + s->qy[i] * s->qy[i] + s->qz[i] * s->qz[i]);
        s->qw[i] = s->qw[i] / qm; s->qx[i] = s->qx[i] / qm;
        s->qy[i] = s->qy[i] / qm; s->qz[i] = s->qz[i] / qm;
    }
}

static void lgv_final(or_sys *s, double dt)
{
    double RM[9];
    s->kin_eng = 0.0;
    s->kin_rot = 0.0;
    for (int i = 0; i < s->n; ++i) {
        half_kick(s, i, dt);
        s->lx[i] = s->lx[i] + 0.5 * dt * s->tx[i];
        s->ly[i] = s->ly[i] + 0.5 * dt * s->ty[i];
        s->lz[i] = s->lz[i] + 0.5 * dt * s->tz[i];
        double kr = kin_rot_of(s, i, RM);
        s->kin_eng = s->kin_eng + 0.5 * s->mass[i] * (s->vx[i] * s->vx[i] + s->vy[i] * s->vy[i] + s->vz[i] * s->vz[i]);
        s->kin_rot = s->kin_rot + kr;
    }
}

void or_lgv_initial_step(or_sys *s, double dt, int step, double damping) { lgv_initial(s, dt, step, damping, 0); }
void or_lgv_final_step(or_sys *s, double dt, double damping) { (void)damping; lgv_final(s, dt); }
void or_lgvc_initial_step(or_sys *s, double dt, int step, double damping)
{
    lgv_initial(s, dt, step, damping, 1);
    or_apply_constraints_a(s, dt);
}
void or_lgvc_final_step(or_sys *s, double dt, double damping)
{
    (void)damping;
    lgv_final(s, dt);
    or_apply_constraints_b(s, dt);
}

/* ---- RATTLE (system.f90:1060-1209) ---- */
int or_apply_constraints_a(or_sys *s, double dt)
{
    const double r_tol = 1.0e-5;
    const int max_rattle = 100;
    int n = s->n, rattle = 0;
    unsigned char *moved = (unsigned char *)malloc((size_t)(n > 0 ? n : 1));
    memset(moved, 1, (size_t)n);
    int status = 0;
    for (;;) {
        int any = 0;
        for (int i = 0; i < n; ++i) if (moved[i]) { any = 1; break; }
        if (!(any && rattle < max_rattle)) break;
        for (int ix = 0; ix < s->nc; ++ix) {
            int I = s->bbi[ix], J = s->bbj[ix];
            double len = s->bb_len[ix];
            if (!(moved[I] || moved[J])) continue;
            double R[3] = {s->rx[I] - s->rx[J], s->ry[I] - s->ry[J], s->rz[I] - s->rz[J]};
            if (s->periodic) for (int k = 0; k < 3; ++k) R[k] = mimg(R[k], s->box[k]);
            double chi = len * len - (R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
            if (fabs(chi) > 2 * r_tol * (len * len)) {
                double Ro[3] = {s->rx_old[I] - s->rx_old[J], s->ry_old[I] - s->ry_old[J], s->rz_old[I] - s->rz_old[J]};
                if (s->periodic) for (int k = 0; k < 3; ++k) Ro[k] = mimg(Ro[k], s->box[k]);
                double vdot = R[0] * Ro[0] + R[1] * Ro[1] + R[2] * Ro[2];
                if (vdot < r_tol * (len * len)) { status = -(1 + ix); goto done; }
                double L = chi / (2.0 * vdot * (s->inv_mass[I] + s->inv_mass[J]));
                s->virial_bonded = s->virial_bonded + (L * (len * len)) / (dt * dt);
                double ai = L * s->inv_mass[I], aj = L * s->inv_mass[J];
                s->rx[I] = s->rx[I] + ai * Ro[0]; s->ry[I] = s->ry[I] + ai * Ro[1]; s->rz[I] = s->rz[I] + ai * Ro[2];
                s->rx[J] = s->rx[J] - aj * Ro[0]; s->ry[J] = s->ry[J] - aj * Ro[1]; s->rz[J] = s->rz[J] - aj * Ro[2];
                s->vx[I] = s->vx[I] + ai * Ro[0] / dt; s->vy[I] = s->vy[I] + ai * Ro[1] / dt; s->vz[I] = s->vz[I] + ai * Ro[2] / dt;
                s->vx[J] = s->vx[J] - aj * Ro[0] / dt; s->vy[J] = s->vy[J] - aj * Ro[1] / dt; s->vz[J] = s->vz[J] - aj * Ro[2] / dt;
                moved[I] = 1; moved[J] = 1;
            } else {
                moved[I] = 0; moved[J] = 0;
            }
        }
        rattle++;
    }
done:
    free(moved);
    return status;
}

int or_apply_constraints_b(or_sys *s, double dt)
{
    const double r_tol = 1.0e-5;
    const int max_rattle = 100;
    int n = s->n, rattle = 0;
    unsigned char *moved = (unsigned char *)malloc((size_t)(n > 0 ? n : 1));
    memset(moved, 1, (size_t)n);
    for (;;) {
        int any = 0;
        for (int i = 0; i < n; ++i) if (moved[i]) { any = 1; break; }
        if (!(any && rattle < max_rattle)) break;
        for (int ix = 0; ix < s->nc; ++ix) {
            int I = s->bbi[ix], J = s->bbj[ix];
            double len = s->bb_len[ix];
            if (!(moved[I] || moved[J])) continue;
            double R[3] = {s->rx[I] - s->rx[J], s->ry[I] - s->ry[J], s->rz[I] - s->rz[J]};
            if (s->periodic) for (int k = 0; k < 3; ++k) R[k] = mimg(R[k], s->box[k]);
            double V[3] = {s->vx[I] - s->vx[J], s->vy[I] - s->vy[J], s->vz[I] - s->vz[J]};
            double vdot = R[0] * V[0] + R[1] * V[1] + R[2] * V[2];
            double L = -vdot / ((s->inv_mass[I] + s->inv_mass[J]) * (len * len));
            if (fabs(L) > r_tol) {
                s->virial_bonded = s->virial_bonded + (L * (len * len)) / dt;
                double ai = L * s->inv_mass[I], aj = L * s->inv_mass[J];
                s->vx[I] = s->vx[I] + ai * R[0]; s->vy[I] = s->vy[I] + ai * R[1]; s->vz[I] = s->vz[I] + ai * R[2];
                s->vx[J] = s->vx[J] - aj * R[0]; s->vy[J] = s->vy[J] - aj * R[1]; s->vz[J] = s->vz[J] - aj * R[2];
                moved[I] = 1; moved[J] = 1;
            } else {
                moved[I] = 0; moved[J] = 0;
            }
        }
        rattle++;
    }
    free(moved);
    return 0;
}

/* ------------------------------------------------------------------ */
/* README pair loop (README.md:86-103) over the half list               */
/* ------------------------------------------------------------------ */
/* one pair (i, j) of kind's routine; returns 1 for the kinds with torques,
   -1 for an unknown kind */
static int pair_terms(const or_sys *s, int i, int j, int kind, const double *p, double *pe_out, double f[3],
                      double ti[3], double tj[3])
{
    double pe = 0.0;
    int torque = 0;
    f[0] = f[1] = f[2] = 0.0;
    ti[0] = ti[1] = ti[2] = 0.0;
    tj[0] = tj[1] = tj[2] = 0.0;
    switch (kind) {
            case 1: or_lj(s, i, j, p[0], p[1], p[2], &pe, f); break;
            case 2: or_lj12(s, i, j, p[0], p[1], p[2], &pe, f); break;
            case 3: or_ljw(s, i, j, p[0], p[1], p[2], p[3], &pe, f); break;
            case 4: or_wca(s, i, j, p[0], p[1], &pe, f); break;
            case 5: or_morse(s, i, j, p[0], p[1], p[2], p[3], &pe, f); break;
            case 6: or_gb(s, i, j, 0, p[0], (int)p[1], (int)p[2], &pe, f, ti, tj); torque = 1; break;
            case 7: or_gb(s, i, j, 1, 0.0, (int)p[1], (int)p[2], &pe, f, ti, tj); torque = 1; break;
            case 8: or_ecp(s, i, j, 0, p[0], &pe, f, ti, tj, NULL, NULL); torque = 1; break;
            case 9: or_ecp(s, i, j, 1, 0.0, &pe, f, ti, tj, NULL, NULL); torque = 1; break;
            case 10: or_dh(s, i, j, p[0], p[1], p[2], &pe, f); break;
            case 11: or_gbx(s, i, j, 2, p[0], p[3], 0.0, (int)p[1], (int)p[2], &pe, f, ti, tj); torque = 1; break;
            case 12: or_gbx(s, i, j, 3, p[0], 0.0, p[3], (int)p[1], (int)p[2], &pe, f, ti, tj); torque = 1; break;
            case 13: or_gbx(s, i, j, 4, p[0], p[3], p[4], (int)p[1], (int)p[2], &pe, f, ti, tj); torque = 1; break;
            case 14: ecp_eval(s, i, j, 2, p[0], p[1], 0.0, -1.0, &pe, f, ti, tj, NULL, NULL); torque = 1; break;
            case 15: ecp_eval(s, i, j, 3, p[0], 0.0, p[1], -1.0, &pe, f, ti, tj, NULL, NULL); torque = 1; break;
            case 16: ecp_eval(s, i, j, 4, p[0], p[1], p[2], -1.0, &pe, f, ti, tj, NULL, NULL); torque = 1; break;
            case 18: or_coulomb(s, i, j, p[0], p[1], &pe, f); break;
            default: return -1;
            }
    *pe_out = pe;
    return torque;
}

/* the caller's pair loop (README.md:86-103) over the half list */
void or_pair_loop(or_sys *s, const or_nlist *nl, int kind, const double *p)
{
    for (int i = 0; i < s->n - 1; ++i) {
        for (long k = nl->offset[i]; k < nl->offset[i + 1]; ++k) {
            int j = nl->nbr[k];
            double pe, f[3], ti[3], tj[3];
            const int torque = pair_terms(s, i, j, kind, p, &pe, f, ti, tj);
            if (torque < 0) return;
            s->pot_eng = s->pot_eng + pe;
            s->fx[i] = s->fx[i] + f[0]; s->fy[i] = s->fy[i] + f[1]; s->fz[i] = s->fz[i] + f[2];
            s->fx[j] = s->fx[j] - f[0]; s->fy[j] = s->fy[j] - f[1]; s->fz[j] = s->fz[j] - f[2];
            if (torque) {
                s->tx[i] = s->tx[i] + ti[0]; s->ty[i] = s->ty[i] + ti[1]; s->tz[i] = s->tz[i] + ti[2];
                s->tx[j] = s->tx[j] + tj[0]; s->ty[j] = s->ty[j] + tj[1]; s->tz[j] = s->tz[j] + tj[2];
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* all-cores variants for the CPU baseline (built with -fopenmp into    */
/* liboracle_omp.so; without OpenMP they run on one thread)             */
/* ------------------------------------------------------------------ */
/* or_pair_loop with the pair terms evaluated by all threads into buf
   (10 doubles per half-list entry, buf sized 10 nnz) and accumulated in the
   loop's order afterwards: the same bits as or_pair_loop */
void or_pair_loop_threads(or_sys *s, const or_nlist *nl, int kind, const double *p, double *buf)
{
    const int n = s->n;
    int torque = 0;
    /* every row's entries are independent: rows are dealt to threads */
#pragma omp parallel for schedule(dynamic, 64) reduction(| : torque)
    for (int i = 0; i < n - 1; ++i) {
        for (long k = nl->offset[i]; k < nl->offset[i + 1]; ++k) {
            double *r = buf + 10 * k;
            int t = pair_terms(s, i, nl->nbr[k], kind, p, &r[0], &r[1], &r[4], &r[7]);
            if (t > 0) torque = 1;
        }
    }
    for (int i = 0; i < n - 1; ++i) {
        for (long k = nl->offset[i]; k < nl->offset[i + 1]; ++k) {
            const int j = nl->nbr[k];
            const double *r = buf + 10 * k;
            s->pot_eng = s->pot_eng + r[0];
            s->fx[i] = s->fx[i] + r[1]; s->fy[i] = s->fy[i] + r[2]; s->fz[i] = s->fz[i] + r[3];
            s->fx[j] = s->fx[j] - r[1]; s->fy[j] = s->fy[j] - r[2]; s->fz[j] = s->fz[j] - r[3];
            if (torque) {
                s->tx[i] = s->tx[i] + r[4]; s->ty[i] = s->ty[i] + r[5]; s->tz[i] = s->tz[i] + r[6];
                s->tx[j] = s->tx[j] + r[7]; s->ty[j] = s->ty[j] + r[8]; s->tz[j] = s->tz[j] + r[9];
            }
        }
    }
}

/* or_nlist_build with the rows counted and filled by all threads (row counts
   in cnt[n]): the same list */
int or_nlist_build_threads(or_sys *s, or_nlist *nl, long *cnt)
{
    double rskin = nl->cutoff + nl->skin;
    nl->r_skin_sq = rskin * rskin;
    const int n = s->n;
    for (int pass = 0; pass < 2; ++pass) {
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < n - 1; ++i) {
            long K = pass ? nl->offset[i] : 0;
            for (int j = i + 1; j < n; ++j) {
                double d0 = s->rx[i] - s->rx[j], d1 = s->ry[i] - s->ry[j], d2 = s->rz[i] - s->rz[j];
                if (s->periodic) {
                    d0 = mimg(d0, s->box[0]); d1 = mimg(d1, s->box[1]); d2 = mimg(d2, s->box[2]);
                }
                double rsq = d0 * d0 + d1 * d1 + d2 * d2;
                if (rsq < nl->r_skin_sq) {
                    if (pass) nl->nbr[K] = j;
                    ++K;
                }
            }
            if (!pass) cnt[i] = K;
        }
        if (!pass) {
            long K = 0;
            for (int i = 0; i < n - 1; ++i) {
                nl->offset[i] = K;
                K += cnt[i];
            }
            if (K > nl->cap) return -1;
            if (n >= 1) nl->offset[n - 1] = K;
            nl->offset[n] = K;
            nl->nnz = K;
        }
    }
    memcpy(nl->x_last, s->rx, sizeof(double) * n);
    memcpy(nl->y_last, s->ry, sizeof(double) * n);
    memcpy(nl->z_last, s->rz, sizeof(double) * n);
    return 0;
}

int or_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
