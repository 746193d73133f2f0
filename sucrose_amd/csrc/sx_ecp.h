// ECP (Perram-Wertheim ellipsoid contact) device code, FP64 (force.f90:1034-1551).
//
// Per-particle quantities that the reference recomputes for every pair (the
// rotation U = R(q)^T, AR = U^T S^2 U, AE = U^T E U, MM = U^T S and sigma_t,
// force.f90:1427-1483) are computed once per step with the same expressions,
// so each pair sees the reference's values bit-for-bit.
#pragma once
#include "sx_pairs.h"

#include <climits>

// per-particle record fields (stored per particle, PP_STRIDE doubles each)
enum {
    PP_AR = 0,     // 6: a11 a21 a31 a22 a32 a33
    PP_AE = 6,     // 6
    PP_MM = 12,    // 9: column-major MM(:,1) MM(:,2) MM(:,3)
    PP_ST = 21,    // sigma_t of the particle
    PP_AMAX = 22,  // largest semi-axis
    PP_N = 23
};

struct Sym3 {
    double a11, a21, a31, a22, a32, a33;
};

// records are stored per particle (array of structures, PP_STRIDE doubles):
// element q of particle i lives at pp[i * PP_STRIDE + q]; `ld` is unused
#define PP_STRIDE 24
#define PPX(pp, q, i) (pp)[(long)(i) * PP_STRIDE + (q)]

// FP32 screen record of a particle: one 64-byte line of 16 floats, the
// position (x, y, z, 0), the AR lower triangle, the largest and the smallest
// semi-axis, and 4 unused
#define ECP_SCR_STRIDE 16

// per-particle records (force.f90:1427-1483 for one particle): FP64 record
// and FP32 screen record
__device__ __forceinline__ void ecp_pp_particle(const DState &s, int i, double *pp, float *scr)
{
    double w = s.f[SX_QW][i], x = s.f[SX_QX][i], y = s.f[SX_QY][i], z = s.f[SX_QZ][i];
    double U[9];
#define Ux(i_, j_) U[((i_) - 1) + 3 * ((j_) - 1)]
    Ux(1, 1) = w * w + x * x - y * y - z * z;
    Ux(1, 2) = 2.0 * (x * y + w * z);
    Ux(1, 3) = 2.0 * (x * z - w * y);
    Ux(2, 1) = 2.0 * (x * y - w * z);
    Ux(2, 2) = w * w - x * x + y * y - z * z;
    Ux(2, 3) = 2.0 * (y * z + w * x);
    Ux(3, 1) = 2.0 * (x * z + w * y);
    Ux(3, 2) = 2.0 * (y * z - w * x);
    Ux(3, 3) = w * w - x * x - y * y + z * z;
    double sx = s.f[SX_SIG_X][i], sy = s.f[SX_SIG_Y][i], sz = s.f[SX_SIG_Z][i];
    double ex = s.f[SX_EPS_X][i], ey = s.f[SX_EPS_Y][i], ez = s.f[SX_EPS_Z][i];
    double e0 = fmax(fmax(ex, ey), ez);
    double S[3] = {sx / 2.0, sy / 2.0, sz / 2.0};
    double B[3] = {S[0] * S[0], S[1] * S[1], S[2] * S[2]};
    // E = ((e0/eps)**(1/2))/4  (force.f90:1452-1454)
    double E[3] = {sqrt(e0 / ex) / 4.0, sqrt(e0 / ey) / 4.0, sqrt(e0 / ez) / 4.0};
    double *rec = pp + (long)i * PP_STRIDE;
    // matmul_atba (xmath.f90:262-279): the reference computes out(r, c) for
    // r <= c; the record keeps (c, r) = (r, c) in lower-triangle order
    const int ij[6][2] = {{1, 1}, {2, 1}, {3, 1}, {2, 2}, {3, 2}, {3, 3}};
    double ar_l[6];
    for (int t = 0; t < 6; ++t) {
        int r_ = ij[t][1], c_ = ij[t][0];
        double ar = Ux(1, r_) * B[0] * Ux(1, c_) + Ux(2, r_) * B[1] * Ux(2, c_) + Ux(3, r_) * B[2] * Ux(3, c_);
        double ae = Ux(1, r_) * E[0] * Ux(1, c_) + Ux(2, r_) * E[1] * Ux(2, c_) + Ux(3, r_) * E[2] * Ux(3, c_);
        rec[PP_AR + t] = ar;
        rec[PP_AE + t] = ae;
        ar_l[t] = ar;
    }
    // MM = U^T . S: column c is row c of U scaled by S(c)
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 3; ++k) rec[PP_MM + 3 * c + k] = Ux(c + 1, k + 1) * S[c];
    rec[PP_ST] = (sx * sy + sz * sz) * sqrt(2.0 * sx * sy) / 8.0;
    double amax = fmax(fmax(S[0], S[1]), S[2]), amin = fmin(fmin(S[0], S[1]), S[2]);
    rec[PP_AMAX] = amax;
#undef Ux
    float4 *s4 = reinterpret_cast<float4 *>(scr + (long)i * ECP_SCR_STRIDE);
    s4[0] = make_float4((float)s.f[SX_RX][i], (float)s.f[SX_RY][i], (float)s.f[SX_RZ][i], 0.0f);
    s4[1] = make_float4((float)ar_l[0], (float)ar_l[1], (float)ar_l[2], (float)ar_l[3]);
    s4[2] = make_float4((float)ar_l[4], (float)ar_l[5], (float)amax, (float)amin);
}

__device__ __forceinline__ Sym3 load_sym(const double *pp, long ld, int base, int i)
{
    (void)ld;
    const double *r = pp + (long)i * PP_STRIDE + base;
    Sym3 m;
    m.a11 = r[0]; m.a21 = r[1]; m.a31 = r[2];
    m.a22 = r[3]; m.a32 = r[4]; m.a33 = r[5];
    return m;
}

// full 3x3 column-major from a symmetric record
__device__ __forceinline__ void sym_full(const Sym3 &m, double o[9])
{
    o[0] = m.a11; o[1] = m.a21; o[2] = m.a31;
    o[3] = m.a21; o[4] = m.a22; o[5] = m.a32;
    o[6] = m.a31; o[7] = m.a32; o[8] = m.a33;
}

// optim_dist (force.f90:1188-1200)
__device__ __forceinline__ double ecp_optim_dist(double l, const Sym3 &A, const Sym3 &B, const double r[3])
{
    double km[3];
    sx_chol_solve((1 - l) * A.a11 + l * B.a11, (1 - l) * A.a21 + l * B.a21, (1 - l) * A.a31 + l * B.a31,
                  (1 - l) * A.a22 + l * B.a22, (1 - l) * A.a32 + l * B.a32, (1 - l) * A.a33 + l * B.a33, r, km);
    return -l * (1 - l) * (r[0] * km[0] + r[1] * km[1] + r[2] * km[2]);
}

// The half list is cut into ECP_NSEG contiguous segments (one per block);
// the screen compacts the survivors of each segment in order into the same
// segment of the work array and records their count (segcount).  No atomics:
// the work list is deterministic and later stages walk the segments.
#define ECP_NSEG 1024
// the apply stage's slots: per particle one 128-byte line of 16 int2, the
// count in the first (.x) and up to ECP_SLOTS (partner, half pair) entries
// after it (a particle with more walks its neighbour row)
#define ECP_SLOT_LINE 16
#define ECP_SLOTS (ECP_SLOT_LINE - 1)

__device__ __forceinline__ void ecp_segment(long nnz, int b, long &start, long &end)
{
    long chunk = (nnz + ECP_NSEG - 1) / ECP_NSEG;
    start = (long)b * chunk;
    end = start + chunk < nnz ? start + chunk : nnz;
    if (start > nnz) start = nnz;
}

#define ECP_PSI 0.3819660112501051  /* 0.5*(3-sqrt(5)) in FP64, force.f90:1050 */

// brent (force.f90:1046-1172) continuing from the first evaluation fc = f(psi);
// atol/rtol are the reference's 1e-3/1e-2 unless a caller tightens them
__device__ __forceinline__ double ecp_brent(const Sym3 &A, const Sym3 &B, const double r[3], double fc, int &iters,
                                            double atol = 1e-3, double rtol = 1e-2)
{
    const double psi = ECP_PSI;
    double a = 0.0, b = 1.0, deltax = 0.0, rat = 0.0;
    double c = psi;
    double v = c, w = v, x = w;
    double fv = fc, fw = fv, ft = fw;
    int iter = 0;
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
        double fu = ecp_optim_dist(u, A, B, r);
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
    iters = iter;
    return x;
}

// F(l) = l(1-l) r^T G^-1 r and its first two derivatives (G = (1-l)A + lB)
__device__ __forceinline__ void ecp_F_derivs(double l, const Sym3 &A, const Sym3 &B, const double r[3], double &F,
                                             double &d1, double &d2)
{
    double g11 = (1 - l) * A.a11 + l * B.a11, g21 = (1 - l) * A.a21 + l * B.a21, g31 = (1 - l) * A.a31 + l * B.a31;
    double g22 = (1 - l) * A.a22 + l * B.a22, g32 = (1 - l) * A.a32 + l * B.a32, g33 = (1 - l) * A.a33 + l * B.a33;
    double k[3];
    sx_chol_solve(g11, g21, g31, g22, g32, g33, r, k);
    double d11 = B.a11 - A.a11, d21 = B.a21 - A.a21, d31 = B.a31 - A.a31;
    double d22 = B.a22 - A.a22, d32 = B.a32 - A.a32, d33 = B.a33 - A.a33;
    double u[3] = {d11 * k[0] + d21 * k[1] + d31 * k[2], d21 * k[0] + d22 * k[1] + d32 * k[2],
                   d31 * k[0] + d32 * k[1] + d33 * k[2]};
    double w[3];
    sx_chol_solve(g11, g21, g31, g22, g32, g33, u, w);
    double rk = r[0] * k[0] + r[1] * k[1] + r[2] * k[2];
    double kdk = k[0] * u[0] + k[1] * u[1] + k[2] * u[2];
    double udw = u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
    double ll = l * (1 - l);
    F = ll * rk;
    d1 = (1 - 2 * l) * rk - ll * kdk;
    d2 = -2 * rk - 2 * (1 - 2 * l) * kdk + 2 * ll * udw;
}

// maximiser of F on (0,1): Newton on F' = 0 from l0.  F is concave on (0,1)
// with F'(0) > 0 > F'(1), so the sign of F' keeps a bracket [lo, hi]; a Newton
// step below tol ends the solve (that step is taken), a step that leaves the
// bracket is replaced by bisection.  `ok` is 0 only if 100 iterations pass.
__device__ __forceinline__ double ecp_lambda_newton(const Sym3 &A, const Sym3 &B, const double r[3], double l0,
                                                    double tol, int &iters, int &ok, const double *pre = nullptr)
{
    // pre (may be NULL): F, F', F'' already evaluated at l0 (inside (0, 1))
    double lo = 0.0, hi = 1.0, l = pre ? l0 : fmin(fmax(l0, 1e-6), 1.0 - 1e-6);
    ok = 0;
    int it = 0;
    for (; it < 100; ++it) {
        double F, d1, d2;
        if (it == 0 && pre) {
            F = pre[0]; d1 = pre[1]; d2 = pre[2];
        } else {
            ecp_F_derivs(l, A, B, r, F, d1, d2);
        }
        if (d1 > 0) lo = l; else hi = l;
        if (d2 < 0) {
            double step = -d1 / d2;
            if (fabs(step) < tol) { l = l + step; ++it; ok = 1; break; }
            double ln = l + step;
            l = (ln > lo && ln < hi) ? ln : 0.5 * (lo + hi);
        } else {
            l = 0.5 * (lo + hi);
        }
        if (hi - lo < tol) { ++it; ok = 1; break; }
    }
    iters = it;
    return l;
}

// F(l) <= F(lambda*) for every l (lambda* maximises F), so an F(l) beyond the
// contact threshold proves that the pair gives zero at the converged lambda*:
// ecpr is non-zero only for r <= sigma = r / sqrt(F*), i.e. F* <= 1; ecp only
// for r < sigma + cutoff - sigma0, i.e. F* < (r / (r - cutoff + 1))^2.
__device__ __forceinline__ bool ecp_zero_proven(double F, double rsq, int variant, double cutoff)
{
    double thr = 1.0;
    if (variant == 0) {
        double r = sqrt(rsq);
        double den = r - cutoff + 1.0;
        if (den <= 0.0) return false;
        thr = (r / den) * (r / den);
    }
    return F > thr * (1.0 + 1e-12);
}

// The force/torque evaluation of ecp/ecpr (force.f90:1487-1550) for a given
// lambda; RIJ already minimum-imaged.
// parameters of the ECP family: ecp_ (variant 0), ecpr_ (1), ecpw_ (2, weld
// width wf), ecpc_ (3, chirality), ecpwc_ (4, both)
struct EcpPar {
    double cutoff, wf, chir;
};

template <int VARIANT>
__device__ __forceinline__ void ecp_forces(const DState &s, const double *pp, long ld, int a, int b, const EcpPar &P,
                                           const double RIJ[3], double lambda_R, PairOut &o)
{
    constexpr bool welded = VARIANT == 2 || VARIANT == 4, chiral = VARIANT == 3 || VARIANT == 4;
    const double cutoff = P.cutoff, wf = P.wf;
    const int meu = 2, neu = 1;
    const double sigma_0 = 1.0;
    o.pe = 0.0;
    for (int k = 0; k < 3; ++k) { o.f[k] = 0.0; o.ti[k] = 0.0; o.tj[k] = 0.0; }
    double rij_sq = RIJ[0] * RIJ[0] + RIJ[1] * RIJ[1] + RIJ[2] * RIJ[2];
    double rij_mag = sqrt(rij_sq);
    double rhat[3] = {RIJ[0] / rij_mag, RIJ[1] / rij_mag, RIJ[2] / rij_mag};
    Sym3 AR = load_sym(pp, ld, PP_AR, a), BR = load_sym(pp, ld, PP_AR, b);
    Sym3 AE = load_sym(pp, ld, PP_AE, a), BE = load_sym(pp, ld, PP_AE, b);
    double lambda_E = lambda_R;
    double KE[3], KR[3];
    sx_chol_solve((1 - lambda_E) * AE.a11 + lambda_E * BE.a11, (1 - lambda_E) * AE.a21 + lambda_E * BE.a21,
                  (1 - lambda_E) * AE.a31 + lambda_E * BE.a31, (1 - lambda_E) * AE.a22 + lambda_E * BE.a22,
                  (1 - lambda_E) * AE.a32 + lambda_E * BE.a32, (1 - lambda_E) * AE.a33 + lambda_E * BE.a33, RIJ, KE);
    double eps2 = lambda_E * (1.0 - lambda_E) * (rhat[0] * KE[0] + rhat[1] * KE[1] + rhat[2] * KE[2]) / rij_mag;
    sx_chol_solve((1 - lambda_R) * AR.a11 + lambda_R * BR.a11, (1 - lambda_R) * AR.a21 + lambda_R * BR.a21,
                  (1 - lambda_R) * AR.a31 + lambda_R * BR.a31, (1 - lambda_R) * AR.a22 + lambda_R * BR.a22,
                  (1 - lambda_R) * AR.a32 + lambda_R * BR.a32, (1 - lambda_R) * AR.a33 + lambda_R * BR.a33, RIJ, KR);
    double phi = lambda_R * (1.0 - lambda_R) * (rhat[0] * KR[0] + rhat[1] * KR[1] + rhat[2] * KR[2]) / rij_mag;
    double sigma = 1.0 / sqrt(phi);
    double efac = 2.0 * lambda_E * (1.0 - lambda_E) / rij_sq;
    double rfac = 2.0 * lambda_R * (1.0 - lambda_R) / rij_sq;
    double sigma_0_inv = 1.0 / sigma_0;
    double p32 = sigma * sigma * sigma;
    double switch_ = 0, factor = 0, sx = 0, sdx = 0, rlim;
    if (VARIANT != 1) {
        rlim = welded ? (sigma + cutoff + wf - sigma_0) : (sigma + cutoff - sigma_0);
        switch_ = 0.9 * rlim;
        factor = 1.0 / (rlim - switch_);
        double xx = (rij_mag - switch_) * factor;
        sx = 1.0 - 6.0 * xx * xx * xx * xx * xx + 15.0 * xx * xx * xx * xx - 10.0 * xx * xx * xx;
        sdx = -30.0 * xx * xx * xx * xx + 60.0 * xx * xx * xx - 30.0 * xx * xx;
    } else {
        rlim = sigma;
    }
    bool inside = (VARIANT == 1) ? (rij_mag <= rlim) : (rij_mag < rlim);
    if (!inside) return;
    // eps1 = sigma_t sqrt(det(inverse(AR + BR)))  (force.f90:1482-1488)
    double EAB[9], EM[9];
    {
        double ar[9], br[9];
        sym_full(AR, ar);
        sym_full(BR, br);
        for (int k = 0; k < 9; ++k) EAB[k] = ar[k] + br[k];
#define Mx(m, i, j) (m)[((i) - 1) + 3 * ((j) - 1)]
        Mx(EM, 1, 1) = Mx(EAB, 2, 2) * Mx(EAB, 3, 3) - Mx(EAB, 3, 2) * Mx(EAB, 2, 3);
        Mx(EM, 1, 2) = Mx(EAB, 1, 3) * Mx(EAB, 3, 2) - Mx(EAB, 1, 2) * Mx(EAB, 3, 3);
        Mx(EM, 1, 3) = Mx(EAB, 1, 2) * Mx(EAB, 2, 3) - Mx(EAB, 1, 3) * Mx(EAB, 2, 2);
        Mx(EM, 2, 1) = Mx(EAB, 2, 3) * Mx(EAB, 3, 1) - Mx(EAB, 2, 1) * Mx(EAB, 3, 3);
        Mx(EM, 2, 2) = Mx(EAB, 1, 1) * Mx(EAB, 3, 3) - Mx(EAB, 1, 3) * Mx(EAB, 3, 1);
        Mx(EM, 2, 3) = Mx(EAB, 1, 3) * Mx(EAB, 2, 1) - Mx(EAB, 1, 1) * Mx(EAB, 2, 3);
        Mx(EM, 3, 1) = Mx(EAB, 2, 1) * Mx(EAB, 3, 2) - Mx(EAB, 2, 2) * Mx(EAB, 3, 1);
        Mx(EM, 3, 2) = Mx(EAB, 1, 2) * Mx(EAB, 3, 1) - Mx(EAB, 1, 1) * Mx(EAB, 3, 2);
        Mx(EM, 3, 3) = Mx(EAB, 1, 1) * Mx(EAB, 2, 2) - Mx(EAB, 1, 2) * Mx(EAB, 2, 1);
        double det = Mx(EAB, 1, 1) * (Mx(EAB, 2, 2) * Mx(EAB, 3, 3) - Mx(EAB, 3, 2) * Mx(EAB, 2, 3))
                   - Mx(EAB, 1, 2) * (Mx(EAB, 2, 1) * Mx(EAB, 3, 3) - Mx(EAB, 3, 1) * Mx(EAB, 2, 3))
                   + Mx(EAB, 1, 3) * (Mx(EAB, 2, 1) * Mx(EAB, 3, 2) - Mx(EAB, 3, 1) * Mx(EAB, 2, 2));
        for (int k = 0; k < 9; ++k) EM[k] = EM[k] / det;
    }
    double detEM = Mx(EM, 1, 1) * (Mx(EM, 2, 2) * Mx(EM, 3, 3) - Mx(EM, 3, 2) * Mx(EM, 2, 3))
                 - Mx(EM, 1, 2) * (Mx(EM, 2, 1) * Mx(EM, 3, 3) - Mx(EM, 3, 1) * Mx(EM, 2, 3))
                 + Mx(EM, 1, 3) * (Mx(EM, 2, 1) * Mx(EM, 3, 2) - Mx(EM, 3, 1) * Mx(EM, 2, 2));
    double sigma_t = sqrt(PPX(pp, PP_ST, a) * PPX(pp, PP_ST, b));
    double eps1 = sigma_t * sqrt(detEM);

    double sr;
    if (!welded) sr = sigma_0 / (rij_mag - sigma + sigma_0);
    else if (rij_mag < sigma) sr = sigma_0 / (rij_mag - sigma + sigma_0);
    else if (rij_mag >= sigma && rij_mag <= sigma + wf) sr = 1.0;
    else sr = sigma_0 / (rij_mag - sigma + sigma_0 - wf);
    double sr2 = sr * sr, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
    double srx = (sr12 - 2.0 * sr6);
    if (VARIANT == 1) srx = srx + 1.0;
    double srd = 12.0 * (sr6 - sr12) * sr;
    double e1n = eps1;                 // eps1**neu, neu = 1
    double e2m = eps2 * eps2;          // eps2**meu, meu = 2
    double e2m1 = eps2;                // eps2**(meu-1)
    double dotKR = rhat[0] * KR[0] + rhat[1] * KR[1] + rhat[2] * KR[2];
    double dotKE = rhat[0] * KE[0] + rhat[1] * KE[1] + rhat[2] * KE[2];
    // the chiral forms leave srx / srd out of the derivative terms (force.f90:1889-1914)
    double pref1 = chiral ? e2m * (neu * e1n) : e2m * (neu * e1n) * srx;
    double de1_u1[3], de1_u2[3];
    for (int side = 0; side < 2; ++side) {
        int who = side == 0 ? a : b;
        double d[3][3];
        for (int c = 0; c < 3; ++c) {
            double col[3] = {PPX(pp, PP_MM + 3 * c + 0, who), PPX(pp, PP_MM + 3 * c + 1, who),
                             PPX(pp, PP_MM + 3 * c + 2, who)};
            double t[3];
            t[0] = Mx(EM, 1, 1) * col[0] + Mx(EM, 1, 2) * col[1] + Mx(EM, 1, 3) * col[2];
            t[1] = Mx(EM, 2, 1) * col[0] + Mx(EM, 2, 2) * col[1] + Mx(EM, 2, 3) * col[2];
            t[2] = Mx(EM, 3, 1) * col[0] + Mx(EM, 3, 2) * col[1] + Mx(EM, 3, 3) * col[2];
            cross3(col, t, d[c]);
            for (int k = 0; k < 3; ++k) d[c][k] = pref1 * d[c][k];
        }
        double *dst = side == 0 ? de1_u1 : de1_u2;
        for (int k = 0; k < 3; ++k) dst[k] = (d[0][k] + d[1][k]) + d[2][k];
    }
#undef Mx
    double aef[9], arf[9], bef[9], brf[9];
    sym_full(AE, aef); sym_full(AR, arf); sym_full(BE, bef); sym_full(BR, brf);
    double KE_AE[3], KR_AR[3], KE_BE[3], KR_BR[3];
    // vmdot: a^T . B  (xmath.f90:219-229)
    for (int j = 0; j < 3; ++j) {
        KE_AE[j] = KE[0] * aef[0 + 3 * j] + KE[1] * aef[1 + 3 * j] + KE[2] * aef[2 + 3 * j];
        KR_AR[j] = KR[0] * arf[0 + 3 * j] + KR[1] * arf[1 + 3 * j] + KR[2] * arf[2 + 3 * j];
        KE_BE[j] = KE[0] * bef[0 + 3 * j] + KE[1] * bef[1 + 3 * j] + KE[2] * bef[2 + 3 * j];
        KR_BR[j] = KR[0] * brf[0 + 3 * j] + KR[1] * brf[1 + 3 * j] + KR[2] * brf[2 + 3 * j];
    }
    double ce1[3], cr1[3], ce2[3], cr2[3];
    cross3(KE_AE, KE, ce1);
    cross3(KR_AR, KR, cr1);
    cross3(KE_BE, KE, ce2);
    cross3(KR_BR, KR, cr2);
    double pot = e1n * e2m * srx;
    if (!chiral) {
        for (int k = 0; k < 3; ++k) {
            double dR_dr = sigma_0_inv * (rhat[k] + 0.5 * rfac * (KR[k] - dotKR * rhat[k]) * p32);
            double de2_dr = e1n * (meu * e2m1) * efac * (KE[k] - dotKE * rhat[k]);
            double de2_u1 = e1n * (meu * e2m1) * srx * (1 - lambda_E) * efac * ce1[k];
            double dR_u1 = e1n * e2m * srd * (0.5 * sigma_0_inv) * (1 - lambda_R) * rfac * cr1[k] * p32;
            double de2_u2 = e1n * (meu * e2m1) * srx * lambda_E * efac * ce2[k];
            double dR_u2 = e1n * e2m * srd * (0.5 * sigma_0_inv) * lambda_R * rfac * cr2[k] * p32;
            o.f[k] = (-1.0) * (e1n * e2m * srd * dR_dr + srx * de2_dr);
            o.ti[k] = (de1_u1[k] + de2_u1) + dR_u1;
            o.tj[k] = (de1_u2[k] + de2_u2) + dR_u2;
        }
    } else {
        // ecpc_ / ecpwc_ (force.f90:1889-1968, 2143-2222): the plain terms
        // regrouped, then the chirality-weighted pseudoscalar terms built on
        // the particles' z axes u = R(q) e_z
        double srxc, srdc;
        if (welded) {
            srxc = (sr12 - (12.0 / 7.0) * sr6 * sr);
            srdc = 12.0 * (sr6 * sr2 - sr12 * sr);
        } else {
            srxc = sr6 * sr;
            srdc = -7.0 * sr6 * sr2;
        }
        double u1z[3], u2z[3];
        quat_u(s, a, u1z);
        quat_u(s, b, u2z);
        double u1xu2[3], u2xrij[3], rijxu1[3];
        cross3(u1z, u2z, u1xu2);
        cross3(u2z, rhat, u2xrij);
        cross3(rhat, u1z, rijxu1);
        const double uur = u1xu2[0] * rhat[0] + u1xu2[1] * rhat[1] + u1xu2[2] * rhat[2];
        const double uu = u1z[0] * u2z[0] + u1z[1] * u2z[1] + u1z[2] * u2z[2];
        double term1[3], term2[3], tu1[3], tu2[3];
        for (int k = 0; k < 3; ++k) {
            term1[k] = (e1n * e2m * srxc * uu * u2xrij[k] + e1n * e2m * srxc * uur * u2z[k]);
            term2[k] = (e1n * e2m * srxc * uu * rijxu1[k] + e1n * e2m * srxc * uur * u1z[k]);
        }
        cross3(u1z, term1, tu1);
        cross3(u2z, term2, tu2);
        const double chir = P.chir;
        for (int k = 0; k < 3; ++k) {
            double dR_dr = e1n * e2m * sigma_0_inv * (rhat[k] + 0.5 * rfac * (KR[k] - dotKR * rhat[k]) * p32);
            double de2_dr = e1n * (meu * e2m1) * efac * (KE[k] - dotKE * rhat[k]);
            double de2_u1 = e1n * (meu * e2m1) * (1 - lambda_E) * efac * ce1[k];
            double dR_u1 = e1n * e2m * (0.5 * sigma_0_inv) * (1 - lambda_R) * rfac * cr1[k] * p32;
            double de2_u2 = e1n * (meu * e2m1) * lambda_E * efac * ce2[k];
            double dR_u2 = e1n * e2m * (0.5 * sigma_0_inv) * lambda_R * rfac * cr2[k] * p32;
            double f = (-1.0) * (srx * de2_dr + srd * dR_dr);
            double t1 = (srx * (de1_u1[k] + de2_u1) + srd * dR_u1);
            double t2 = (srx * (de1_u2[k] + de2_u2) + srd * dR_u2);
            double dRr = dR_dr * srdc * uur * uu, de2r = de2_dr * srxc * uur * uu;
            double duur_r = e1n * e2m * srxc * uu * (u1xu2[k] - uur * rhat[k]) / rij_mag;
            double fc = (-1.0) * (de2r + dRr + duur_r);
            double t1c = (de1_u1[k] * srxc * uur * uu + de2_u1 * srxc * uur * uu + dR_u1 * srdc * uur * uu - tu1[k]);
            double t2c = (de1_u2[k] * srxc * uur * uu + de2_u2 * srxc * uur * uu + dR_u2 * srdc * uur * uu - tu2[k]);
            o.f[k] = f + chir * fc;
            o.ti[k] = t1 + chir * t1c;
            o.tj[k] = t2 + chir * t2c;
        }
        pot = pot + chir * (e1n * e2m * srxc * uur * uu);
    }
    if (VARIANT != 1 && (rij_mag >= switch_) && (rij_mag <= rlim)) {
        for (int k = 0; k < 3; ++k) {
            o.f[k] = sx * o.f[k] - sdx * pot * factor;
            o.ti[k] = sx * o.ti[k] - sdx * pot * factor;
            o.tj[k] = sx * o.tj[k] - sdx * pot * factor;
        }
        pot = sx * pot;
    }
    o.pe = pot;
}

// ---------------------------------------------------------------------------
// one pair of the solve stage
// ---------------------------------------------------------------------------
template <int VARIANT>
__device__ __forceinline__ double ecp_solve_lambda(int solver, const double *pp, long ld, int a, int b,
                                                   const double RIJ[3], double fpsi, double lam0, int &iters,
                                                   int &fallback)
{
    Sym3 A = load_sym(pp, ld, PP_AR, a), B = load_sym(pp, ld, PP_AR, b);
    fallback = 0;
    if (solver == SX_ECP_BRENT_REFERENCE) return ecp_brent(A, B, RIJ, fpsi, iters);
    int ok = 0;
    double l = ecp_lambda_newton(A, B, RIJ, lam0, 1e-10, iters, ok);
    if (!ok) {
        int it2 = 0;
        l = ecp_lambda_newton(A, B, RIJ, ECP_PSI, 1e-10, it2, ok);
        iters += it2;
        fallback = 1;
    }
    return l;
}

// what the apply stage reads: the slots, and the list for particles whose
// pairs overflowed them
struct EcpApplyArgs {
    int2 *slot;
    const int *row_start, *nbr, *lower, *half_start;
    const double *res;
};

// the current context's apply-stage arguments
EcpApplyArgs sx_ecp_apply_args(sx_ctx *c);

// compare-and-swap on the partner index
__device__ __forceinline__ void slot_cas(int2 &a, int2 &b)
{
    const bool sw = b.x < a.x;
    const int2 lo = sw ? b : a, hi = sw ? a : b;
    a = lo;
    b = hi;
}

// the first N entries sorted by partner (bitonic network, N a power of two)
template <int N>
__device__ __forceinline__ void slot_sort(int2 (&e)[ECP_SLOT_LINE])
{
#pragma unroll
    for (int k = 2; k <= N; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const int l = i ^ j;
                if (l > i) {
                    if ((i & k) == 0) slot_cas(e[i], e[l]);
                    else slot_cas(e[l], e[i]);
                }
            }
}

// pair k's results added to particle i's sums: i is the pair's first
// particle (force +F, its torque, the energy) or its second (force -F, the
// other torque)
__device__ __forceinline__ void ecp_add(double (&acc)[7], const double *res, long k, bool first)
{
    const double *r = res + k * 10;
    if (first) {
        acc[0] = acc[0] + r[1]; acc[1] = acc[1] + r[2]; acc[2] = acc[2] + r[3];
        acc[3] = acc[3] + r[4]; acc[4] = acc[4] + r[5]; acc[5] = acc[5] + r[6];
        acc[6] = acc[6] + r[0];
    } else {
        acc[0] = acc[0] - r[1]; acc[1] = acc[1] - r[2]; acc[2] = acc[2] - r[3];
        acc[3] = acc[3] + r[7]; acc[4] = acc[4] + r[8]; acc[5] = acc[5] + r[9];
    }
}

// Particle i's sums (force, torque, energy) over its solved pairs, from its
// slots sorted by partner, so that they are added in the order of its
// neighbour row (ascending partner: the order in which the reference's pair
// loop reaches them); pairs the learned solve proved out of range (half pair
// -1) are skipped.  A particle with more than ECP_SLOTS pairs walks its row
// instead: a forward entry's half pair follows from the row's own start, a
// reverse entry's from a binary search of the partner's row, and live[k]
// says whether the pair was solved.  The slot count is cleared for the next
// solve.
__device__ __forceinline__ void ecp_particle_sums(int i, const EcpApplyArgs &a, const unsigned char *live,
                                                  double (&acc)[7])
{
    // the particle's whole line in one round trip
    int4 w[ECP_SLOT_LINE / 2];
    const int4 *line = reinterpret_cast<const int4 *>(a.slot + (long)i * ECP_SLOT_LINE);
#pragma unroll
    for (int u = 0; u < ECP_SLOT_LINE / 2; ++u) w[u] = line[u];
    const int c = w[0].x;
    a.slot[(long)i * ECP_SLOT_LINE].x = 0;
    for (int q = 0; q < 7; ++q) acc[q] = 0.0;
    if (c <= ECP_SLOTS) {
        int2 e[ECP_SLOT_LINE];
#pragma unroll
        for (int u = 0; u < ECP_SLOT_LINE; ++u) {
            const int wi = (u + 1) >> 1;   // entry u is word u + 1 of the line
            const int2 v = (u + 1) & 1 ? make_int2(w[wi].z, w[wi].w) : make_int2(w[wi].x, w[wi].y);
            e[u] = u < c ? v : make_int2(INT_MAX, -1);
        }
        if (c > 8) slot_sort<16>(e);
        else if (c > 4) slot_sort<8>(e);
        else if (c > 1) slot_sort<4>(e);
        // four entries' results requested together (an unused entry reads
        // pair 0), then added in order
#pragma unroll
        for (int b0 = 0; b0 < ECP_SLOT_LINE; b0 += 4) {
            if (b0 >= c) break;
            double r[4][10];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const bool use = b0 + u < c && e[b0 + u].y >= 0;
                const double *p = a.res + (long)(use ? e[b0 + u].y : 0) * 10;
#pragma unroll
                for (int q = 0; q < 10; ++q) r[u][q] = p[q];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (!(b0 + u < c && e[b0 + u].y >= 0)) continue;
                if (e[b0 + u].x > i) {
                    acc[0] = acc[0] + r[u][1]; acc[1] = acc[1] + r[u][2]; acc[2] = acc[2] + r[u][3];
                    acc[3] = acc[3] + r[u][4]; acc[4] = acc[4] + r[u][5]; acc[5] = acc[5] + r[u][6];
                    acc[6] = acc[6] + r[u][0];
                } else {
                    acc[0] = acc[0] - r[u][1]; acc[1] = acc[1] - r[u][2]; acc[2] = acc[2] - r[u][3];
                    acc[3] = acc[3] + r[u][7]; acc[4] = acc[4] + r[u][8]; acc[5] = acc[5] + r[u][9];
                }
            }
        }
    } else {
        const int b = a.row_start[i], e = a.row_start[i + 1], lo = a.lower[i], hs = a.half_start[i];
        for (int q = b; q < e; ++q) {
            const int j = a.nbr[q];
            long k;
            if (j > i) {
                k = hs + (q - b - lo);
            } else {
                const int bj = a.row_start[j], lj = a.lower[j];
                int l = bj + lj, h = a.row_start[j + 1];
                while (l < h) {
                    const int m = (l + h) >> 1;
                    if (a.nbr[m] < i) l = m + 1; else h = m;
                }
                k = a.half_start[j] + (l - bj - lj);
            }
            if (live[k]) ecp_add(acc, a.res, k, j > i);
        }
    }
}

// the apply stage's per-particle slots (always taken: a caller with no apply
// stage after the solve clears the counts itself)
struct EcpSlots {
    int2 *slot;
};

// Solves pair k = (a, b) of the work list and stores its energy, force and
// torques at res[10 k]; the pair takes a slot of a, (b, k), and one of b,
// (a, k), for the apply stage (k = -1 where the warm start proves the pair
// out of range; a count beyond ECP_SLOTS keeps rising without a slot, and
// the apply stage then walks that particle's row).  warm: l0 is a warm start (learned), and a pair whose
// F(l0) already proves it out of range ends there (returns 1, live[k] = 0).
template <int VARIANT>
__device__ __forceinline__ int ecp_pair_solve(const DState &s, const double *pp, long ld, int k, int a, int b,
                                              int solver, bool warm, double l0, const EcpPar &P, double *res,
                                              unsigned char *live, const EcpSlots &sl, int &iters, int &fb)
{
    double RIJ[3];
    pair_rij(s, a, b, RIJ);
    iters = 0;
    fb = 0;
    // the slots are taken once the pair's records are requested (the
    // counts' round trip then overlaps the solve instead of delaying the
    // records) and written at the end
    int sa = ECP_SLOTS, sb = ECP_SLOTS;
    auto take = [&]() {
        sa = atomicAdd(&sl.slot[(long)a * ECP_SLOT_LINE].x, 1);
        sb = atomicAdd(&sl.slot[(long)b * ECP_SLOT_LINE].x, 1);
    };
    auto put = [&](int kk) {
        if (sa < ECP_SLOTS) sl.slot[(long)a * ECP_SLOT_LINE + 1 + sa] = make_int2(b, kk);
        if (sb < ECP_SLOTS) sl.slot[(long)b * ECP_SLOT_LINE + 1 + sb] = make_int2(a, kk);
    };
    double lam;
    if (solver == SX_ECP_BRENT_REFERENCE) {
        Sym3 A = load_sym(pp, ld, PP_AR, a), B = load_sym(pp, ld, PP_AR, b);
        take();
        lam = ecp_brent(A, B, RIJ, ecp_optim_dist(ECP_PSI, A, B, RIJ), iters);
    } else {
        if (warm) {
            // F and its derivatives at the warm start serve both the
            // out-of-range proof and the first Newton step
            Sym3 A = load_sym(pp, ld, PP_AR, a), B = load_sym(pp, ld, PP_AR, b);
            take();
            double rsq = RIJ[0] * RIJ[0] + RIJ[1] * RIJ[1] + RIJ[2] * RIJ[2];
            l0 = fmin(fmax(l0, 1e-6), 1.0 - 1e-6);
            double d[3];
            ecp_F_derivs(l0, A, B, RIJ, d[0], d[1], d[2]);
            if (ecp_zero_proven(d[0], rsq, VARIANT == 1 ? 1 : 0, P.cutoff + P.wf)) {
                live[k] = 0;
                put(-1);
                return 1;
            }
            int ok = 0;
            lam = ecp_lambda_newton(A, B, RIJ, l0, 1e-10, iters, ok, d);
            if (!ok) {
                int it2 = 0;
                lam = ecp_lambda_newton(A, B, RIJ, ECP_PSI, 1e-10, it2, ok);
                iters += it2;
                fb = 1;
            }
        } else {
            take();
            lam = ecp_solve_lambda<VARIANT>(solver, pp, ld, a, b, RIJ, 0.0, ECP_PSI, iters, fb);
        }
    }
    PairOut o;
    ecp_forces<VARIANT>(s, pp, ld, a, b, P, RIJ, lam, o);
    double *r = res + (long)k * 10;
    r[0] = o.pe;
    r[1] = o.f[0]; r[2] = o.f[1]; r[3] = o.f[2];
    r[4] = o.ti[0]; r[5] = o.ti[1]; r[6] = o.ti[2];
    r[7] = o.tj[0]; r[8] = o.tj[1]; r[9] = o.tj[2];
    put(k);
    return 0;
}
