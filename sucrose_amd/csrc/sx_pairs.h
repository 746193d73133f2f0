// Device pair kernels of force.f90, FP64, one canonical evaluation (I = a,
// J = b, RIJ = R(a) - R(b)) per call.  Expressions keep the reference's
// evaluation order (the translation unit is built with -ffp-contract=off), so
// a pair's PE, FXIJ, TXI, TXJ equal the reference's to the last bit for the
// same FP64 inputs, transcendental functions aside.
//
// The FP32 pair mode.  The cores of LJ, WCA and Debye-Hueckel are templates
// on the arithmetic type T: T = double is the form above, T = float the same
// expressions in FP32 with the hardware square root and exp2 (about 1 ulp
// each) and the hardware reciprocal refined by a Newton step.  gb_, too
// ill-conditioned for FP32, has a form of its own with fewer divisions
// (gb_core_f32).  In all of them the separation, the cutoff test that decides
// whether a pair interacts and every sum over pairs stay FP64.
#pragma once
#include "sx_internal.h"

template <typename T>
struct PairOutT {
    T pe;
    T f[3];
    T ti[3];
    T tj[3];
};
using PairOut = PairOutT<double>;

__device__ __forceinline__ double sx_dv(double a, double b) { return a / b; }
// the hardware reciprocal refined by one Newton step (x + x (1 - b x))
__device__ __forceinline__ float sx_dv(float a, float b)
{
    const float x = __builtin_amdgcn_rcpf(b);
    return a * __builtin_fmaf(__builtin_fmaf(-b, x, 1.0f), x, x);
}
__device__ __forceinline__ double sx_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float sx_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double sx_exp(double x) { return exp(x); }
// e^x = 2^(x log2 e)
__device__ __forceinline__ float sx_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }

__device__ __forceinline__ void pair_rij(const DState &s, int a, int b, double r[3])
{
    r[0] = s.f[SX_RX][a] - s.f[SX_RX][b];
    r[1] = s.f[SX_RY][a] - s.f[SX_RY][b];
    r[2] = s.f[SX_RZ][a] - s.f[SX_RZ][b];
    if (s.periodic) {
        r[0] = sx_mimg(r[0], s.box[0]);
        r[1] = sx_mimg(r[1], s.box[1]);
        r[2] = sx_mimg(r[2], s.box[2]);
    }
}

// lj_calculate_forces (force.f90:6-49) on a minimum-image separation r = R(a) - R(b)
template <typename T = double>
__device__ __forceinline__ void lj_core(const double r[3], const double *p, PairOutT<T> &o)
{
    const double rsq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    const bool cmask = rsq < p[2] * p[2];
    T sigma = T(p[0]), epsilon = T(p[1]), cutoff = T(p[2]);
    T rij_sq = T(rsq);
    T rcut_sq = cutoff * cutoff;
    T sr2 = sx_dv(sigma * sigma, rij_sq), sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
    T sc2 = sx_dv(sigma * sigma, rcut_sq), sc6 = sc2 * sc2 * sc2, sc12 = sc6 * sc6;
    T pot = T(4) * epsilon * (sr12 - sr6);
    T pot_cut = T(4) * epsilon * (sc12 - sc6);
    T coeff = T(24) * epsilon * (T(2) * sr12 - sr6);
    o.pe = cmask ? (pot - pot_cut) : T(0);
    for (int k = 0; k < 3; ++k) o.f[k] = cmask ? sx_dv(coeff * T(r[k]), rij_sq) : T(0);
}

__device__ __forceinline__ void d_lj(const DState &s, int a, int b, const double *p, PairOut &o)
{
    double r[3];
    pair_rij(s, a, b, r);
    lj_core(r, p, o);
}

// lj12_calculate_forces (force.f90:51-93)
__device__ __forceinline__ void d_lj12(const DState &s, int a, int b, const double *p, PairOut &o)
{
    double sigma = p[0], epsilon = p[1], cutoff = p[2];
    double r[3];
    pair_rij(s, a, b, r);
    o.pe = 0.0; o.f[0] = o.f[1] = o.f[2] = 0.0;
    double rij_sq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    double rcut_sq = cutoff * cutoff;
    if (rij_sq < rcut_sq) {
        double sr2 = (sigma * sigma) / rij_sq, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
        double sc2 = (sigma * sigma) / rcut_sq, sc6 = sc2 * sc2 * sc2, sc12 = sc6 * sc6;
        double pot = epsilon * (sr12 - 2.0 * sr6);
        double pot_cut = epsilon * (sc12 - 2.0 * sc6);
        double coeff = 12.0 * epsilon * (sr12 - sr6);
        o.pe = pot - pot_cut;
        for (int k = 0; k < 3; ++k) o.f[k] = coeff * r[k] / rij_sq;
    }
}

// ljw_calculate_forces (force.f90:95-146)
__device__ __forceinline__ void d_ljw(const DState &s, int a, int b, const double *p, PairOut &o)
{
    double sigma = p[0], epsilon = p[1], cutoff = p[2], wf = p[3];
    double r[3];
    pair_rij(s, a, b, r);
    o.pe = 0.0; o.f[0] = o.f[1] = o.f[2] = 0.0;
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
        o.pe = pot - pot_cut;
        for (int k = 0; k < 3; ++k) o.f[k] = coeff * r[k] / rij_mag;
    }
}

// wca_calculate_forces (force.f90:155-190)
template <typename T = double>
__device__ __forceinline__ void wca_core(const double r[3], const double *p, PairOutT<T> &o)
{
    o.pe = T(0); o.f[0] = o.f[1] = o.f[2] = T(0);
    const double rsq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    if (rsq <= p[0] * p[0]) {
        T sigma = T(p[0]), epsilon = T(p[1]);
        T rij_sq = T(rsq);
        T sigma_sq = sigma * sigma;
        T sr2 = sx_dv(sigma_sq, rij_sq), sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
        o.pe = epsilon * (sr12 - T(2) * sr6) + epsilon;
        T coeff = T(12) * epsilon * (sr12 - sr6);
        for (int k = 0; k < 3; ++k) o.f[k] = sx_dv(coeff * T(r[k]), rij_sq);
    }
}

__device__ __forceinline__ void d_wca(const DState &s, int a, int b, const double *p, PairOut &o)
{
    double r[3];
    pair_rij(s, a, b, r);
    wca_core(r, p, o);
}

// morse_calculate_forces (force.f90:199-233)
__device__ __forceinline__ void d_morse(const DState &s, int a, int b, const double *p, PairOut &o)
{
    double dissoc = p[0], width = p[1], rmin = p[2], cutoff = p[3];
    double r[3];
    pair_rij(s, a, b, r);
    o.pe = 0.0; o.f[0] = o.f[1] = o.f[2] = 0.0;
    double rij_sq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    double rij_mag = sqrt(rij_sq);
    if (rij_mag < cutoff) {
        double e = exp(-width * (rij_mag - rmin));
        double ec = p[4];   // exp(-width*(cutoff-rmin)), evaluated on the host
        double pot = dissoc * ((1.0 - e) * (1.0 - e) - 1.0);
        double pot_cut = dissoc * ((1.0 - ec) * (1.0 - ec) - 1.0);
        o.pe = pot - pot_cut;
        for (int k = 0; k < 3; ++k) o.f[k] = -2.0 * width * dissoc * (1.0 - e) * e * r[k] / rij_mag;
    }
}

// dh_calculate_forces (force.f90:2589-2617) for charges qa, qb; in FP32 the
// cutoff test compares the FP64 squared distance with the squared cutoff
template <typename T = double>
__device__ __forceinline__ void dh_core(const double r[3], double qa, double qb, const double *p, PairOutT<T> &o)
{
    const T PI_REF = T(3.1415926535897932384626433832795);
    T cutoff = T(p[0]), db_len = T(p[1]), dielectric = T(p[2]);
    const double rsq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    T rij_sq = T(rsq);
    T rij_mag = sx_sqrt(rij_sq);
    bool cmask;
    if constexpr (sizeof(T) == sizeof(double)) cmask = rij_mag < cutoff;
    else cmask = rsq < p[0] * p[0];
    T pe = sx_dv(T(qa) * T(qb) * sx_exp(sx_dv(-rij_mag, db_len)), T(4) * PI_REF * dielectric * rij_mag);
    T fac = pe * (T(1) + sx_dv(rij_mag, db_len));
    o.pe = cmask ? pe : T(0);
    for (int k = 0; k < 3; ++k) o.f[k] = cmask ? sx_dv(fac * T(r[k]), rij_sq) : T(0);
}

__device__ __forceinline__ void d_dh(const DState &s, int a, int b, const double *p, PairOut &o)
{
    double r[3];
    pair_rij(s, a, b, r);
    dh_core(r, s.f[SX_CHARGE][a], s.f[SX_CHARGE][b], p, o);
}

// coulombic_calculate_forces (force.f90:2627-2665): switched off over the
// last 2 units before the cutoff, the switched force subtracting
// sdx * PE * factor (PE already switched) from every component
__device__ __forceinline__ void d_coulomb(const DState &s, int a, int b, const double *p, PairOut &o)
{
    const double PI_REF = 3.1415926535897932384626433832795;
    double cutoff = p[0], dielectric = p[1];
    double r[3];
    pair_rij(s, a, b, r);
    double rij_sq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    double rij_mag = sqrt(rij_sq);
    o.pe = 0.0;
    o.f[0] = o.f[1] = o.f[2] = 0.0;
    if (!(rij_mag < cutoff)) return;
    double sw = cutoff - 2.0;
    double factor = 1.0 / (cutoff - sw);
    double pe = s.f[SX_CHARGE][a] * s.f[SX_CHARGE][b] / (4.0 * PI_REF * dielectric * rij_mag);
    for (int k = 0; k < 3; ++k) o.f[k] = pe * r[k] / rij_sq;
    if (rij_mag >= sw) {
        double xx = (rij_mag - sw) * factor;
        double sx = 1.0 - 6.0 * xx * xx * xx * xx * xx + 15.0 * xx * xx * xx * xx - 10.0 * xx * xx * xx;
        double sdx = -30.0 * xx * xx * xx * xx + 60.0 * xx * xx * xx - 30.0 * xx * xx;
        pe = sx * pe;
        for (int k = 0; k < 3; ++k) o.f[k] = sx * o.f[k] - sdx * pe * factor;
    }
    o.pe = pe;
}

// ---------------------------------------------------------------------------
// Gay-Berne (force.f90:252-533)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T d_ipow(T x, int n)
{
    // binary powering, the lowering of the reference's constant integer powers
    if (n == 0) return T(1);
    T result = T(1), base = x;
    bool first = true;
    while (n > 0) {
        if (n & 1) {
            if (first) { result = base; first = false; }
            else result = result * base;
        }
        n >>= 1;
        if (n) base = base * base;
    }
    return result;
}

struct GbGeo {
    double ru1, ru2, uu, rsq;
    double u1[3], u2[3], r[3];
};

__device__ __forceinline__ double gb_g(double chie, const GbGeo &q)
{
    double t1 = (q.ru1 + q.ru2) * (q.ru1 + q.ru2) / (1 + chie * q.uu);
    double t2 = (q.ru1 - q.ru2) * (q.ru1 - q.ru2) / (1 - chie * q.uu);
    return 1 - (chie / 2.0 / q.rsq) * (t1 + t2);
}

__device__ __forceinline__ void gb_dgdr(double chie, const GbGeo &q, double g, double res[3])
{
    double a = (q.ru1 + q.ru2) / (1 + chie * q.uu);
    double b = (q.ru1 - q.ru2) / (1 - chie * q.uu);
    for (int k = 0; k < 3; ++k) {
        double t1 = a * (q.u1[k] + q.u2[k]);
        double t2 = b * (q.u1[k] - q.u2[k]);
        res[k] = (-chie / q.rsq) * (t1 + t2) + (2 * q.r[k] / q.rsq) * (1 - g);
    }
}

__device__ __forceinline__ void gb_dgdu(double chie, const GbGeo &q, int which, double res[3])
{
    double t1 = (q.ru1 + q.ru2) / (1 + chie * q.uu);
    double t2 = (q.ru1 - q.ru2) / (1 - chie * q.uu);
    for (int k = 0; k < 3; ++k) {
        if (which == 1)
            res[k] = (-chie * q.r[k] / q.rsq) * (t1 + t2) + ((chie * chie * q.u2[k]) / (2 * q.rsq)) * (t1 * t1 - t2 * t2);
        else
            res[k] = (-chie * q.r[k] / q.rsq) * (t1 - t2) + ((chie * chie * q.u1[k]) / (2 * q.rsq)) * (t1 * t1 - t2 * t2);
    }
}

__device__ __forceinline__ void quat_u(const DState &s, int i, double u[3])
{
    double w = s.f[SX_QW][i], x = s.f[SX_QX][i], y = s.f[SX_QY][i], z = s.f[SX_QZ][i];
    u[0] = 2.0 * (x * z + w * y);
    u[1] = 2.0 * (y * z - w * x);
    u[2] = w * w - x * x - y * y + z * z;
}

__device__ __forceinline__ void cross3(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// The per-particle factors of gb_ that depend on particle I alone: its axis
// u = R(q) e_z, chi from its sig, chi' from its eps (eps^(1/mu) for mu != 1)
// and the sig_x / eps_z that scale sigma and epsilon.
struct GbSelf {
    double u[3];
    double chi, xhi, lprp, epsz;
};

__device__ __forceinline__ void gb_self(const DState &s, int a, int mu, GbSelf &g)
{
    quat_u(s, a, g.u);
    double lpar = s.f[SX_SIG_Z][a], lprp = s.f[SX_SIG_X][a];
    double epar = s.f[SX_EPS_Z][a], eprp = s.f[SX_EPS_X][a];
    g.chi = (lpar * lpar - lprp * lprp) / (lpar * lpar + lprp * lprp);
    double ep = eprp, ea = epar;
    if (mu != 1) {
        double muinv = 1.0 / mu;
        ep = pow(eprp, muinv);
        ea = pow(epar, muinv);
    }
    g.xhi = (ep - ea) / (ep + ea);
    g.lprp = lprp;
    g.epsz = epar;
}

// gb_ (variant 0, cutoff + switch), gbr_ (1), gbw_ (2, weld width wf), gbc_
// (3, chirality) and gbwc_ (4, both) for I = a with factors A, J with axis
// u2, RIJ = r
template <int VARIANT>
__device__ __forceinline__ void gb_core(const double r[3], const GbSelf &A, const double u2[3], double cutoff, int mu,
                                        int nu, PairOut &o, double wf = 0.0, double chir = 0.0)
{
    constexpr bool welded = VARIANT == 2 || VARIANT == 4, chiral = VARIANT == 3 || VARIANT == 4;
    const double sigma_0 = 1.0;
    GbGeo q;
    o.pe = 0.0;
    for (int k = 0; k < 3; ++k) { o.f[k] = 0.0; o.ti[k] = 0.0; o.tj[k] = 0.0; }
    for (int k = 0; k < 3; ++k) { q.r[k] = r[k]; q.u1[k] = A.u[k]; q.u2[k] = u2[k]; }
    q.rsq = q.r[0] * q.r[0] + q.r[1] * q.r[1] + q.r[2] * q.r[2];
    double rij_mag = sqrt(q.rsq);
    double rhat[3] = {q.r[0] / rij_mag, q.r[1] / rij_mag, q.r[2] / rij_mag};
    q.ru1 = q.r[0] * q.u1[0] + q.r[1] * q.u1[1] + q.r[2] * q.u1[2];
    q.ru2 = q.r[0] * q.u2[0] + q.r[1] * q.u2[1] + q.r[2] * q.u2[2];
    q.uu = q.u1[0] * q.u2[0] + q.u1[1] * q.u2[1] + q.u1[2] * q.u2[2];
    double chi = A.chi, xhi = A.xhi, lprp = A.lprp;
    double gchi = gb_g(chi, q);
    double gxhi = gb_g(xhi, q);
    double eps1 = 1.0 / sqrt(1 - (chi * q.uu) * (chi * q.uu));
    double eps2 = gxhi / A.epsz;
    double sigma = lprp / sqrt(gchi);
    double sigma_0_inv = 1.0 / sigma_0;
    double g32 = sigma / gchi;
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
    double sr;
    if (!welded) sr = sigma_0 / (rij_mag - sigma + sigma_0);
    else if (rij_mag < sigma) sr = sigma_0 / (rij_mag - sigma + sigma_0);
    else if (rij_mag >= sigma && rij_mag <= sigma + wf) sr = 1.0;
    else sr = sigma_0 / (rij_mag - sigma + sigma_0 - wf);
    double sr2 = sr * sr, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
    double srx = (sr12 - 2.0 * sr6);
    if (VARIANT == 1) srx = srx + 1.0;
    double srd = 12.0 * (sr6 - sr12) * sr;
    double e1n = d_ipow(eps1, nu), e2m = d_ipow(eps2, mu);
    double e2m1 = d_ipow(eps2, mu - 1), e1n2 = d_ipow(eps1, nu + 2);
    double dgr_c[3], dgr_x[3], dgu1_x[3], dgu1_c[3], dgu2_x[3], dgu2_c[3];
    gb_dgdr(chi, q, gchi, dgr_c);
    gb_dgdr(xhi, q, gxhi, dgr_x);
    gb_dgdu(xhi, q, 1, dgu1_x);
    gb_dgdu(chi, q, 1, dgu1_c);
    gb_dgdu(xhi, q, 2, dgu2_x);
    gb_dgdu(chi, q, 2, dgu2_c);
    double TI[3], TJ[3];
    double pot = e1n * e2m * srx;
    if (!chiral) {
        for (int k = 0; k < 3; ++k) {
            double dR_dr = sigma_0_inv * (rhat[k] + 0.5 * dgr_c[k] * g32);
            double de2_dr = e1n * (mu * e2m1) * dgr_x[k];
            double de1_du1 = e2m * (nu * e1n2) * (chi * chi * q.uu * q.u2[k]);
            double de2_du1 = e1n * (mu * e2m1) * dgu1_x[k];
            double dR_du1 = srd * 0.5 * sigma_0_inv * dgu1_c[k] * g32;
            double de1_du2 = e2m * (nu * e1n2) * (chi * chi * q.uu * q.u1[k]);
            double de2_du2 = e1n * (mu * e2m1) * dgu2_x[k];
            double dR_du2 = srd * 0.5 * sigma_0_inv * dgu2_c[k] * g32;
            o.f[k] = (-1.0) * (e1n * e2m * srd * dR_dr + de2_dr * srx);
            TI[k] = (-1.0) * (srx * (de1_du1 + de2_du1) + (e1n * e2m * dR_du1));
            TJ[k] = (-1.0) * (srx * (de1_du2 + de2_du2) + (e1n * e2m * dR_du2));
        }
        cross3(q.u1, TI, o.ti);
        cross3(q.u2, TJ, o.tj);
    } else {
        // gbc_ / gbwc_ (force.f90:751-801, 931-981): the plain terms regrouped
        // (energy factors inside dR, srx / srd outside), then the
        // chirality-weighted pseudoscalar terms
        double srxc, srdc;
        if (welded) {
            srxc = (sr12 - (12.0 / 7.0) * sr6 * sr);
            srdc = 12.0 * (sr6 * sr2 - sr12 * sr);
        } else {
            srxc = sr6 * sr;
            srdc = -7.0 * sr6 * sr2;
        }
        double u1xu2[3], u2xrij[3], rijxu1[3];
        cross3(q.u1, q.u2, u1xu2);
        cross3(q.u2, rhat, u2xrij);
        cross3(rhat, q.u1, rijxu1);
        const double uur = u1xu2[0] * rhat[0] + u1xu2[1] * rhat[1] + u1xu2[2] * rhat[2];
        const double uu = q.uu;
        double FC[3], TIc[3], TJc[3];
        for (int k = 0; k < 3; ++k) {
            double dR_dr = e1n * e2m * sigma_0_inv * (rhat[k] + 0.5 * dgr_c[k] * g32);
            double de2_dr = e1n * (mu * e2m1) * dgr_x[k];
            double de1_du1 = e2m * (nu * e1n2) * (chi * chi * q.uu * q.u2[k]);
            double de2_du1 = e1n * (mu * e2m1) * dgu1_x[k];
            double dR_du1 = e1n * e2m * 0.5 * sigma_0_inv * dgu1_c[k] * g32;
            double de1_du2 = e2m * (nu * e1n2) * (chi * chi * q.uu * q.u1[k]);
            double de2_du2 = e1n * (mu * e2m1) * dgu2_x[k];
            double dR_du2 = e1n * e2m * 0.5 * sigma_0_inv * dgu2_c[k] * g32;
            o.f[k] = (-1.0) * (srx * de2_dr + srd * dR_dr);
            TI[k] = (-1.0) * (srx * (de1_du1 + de2_du1) + srd * dR_du1);
            TJ[k] = (-1.0) * (srx * (de1_du2 + de2_du2) + srd * dR_du2);
            double duur_r = e1n * e2m * srxc * uu * (u1xu2[k] - uur * rhat[k]) / rij_mag;
            FC[k] = (-1.0) * (de2_dr * srxc * uur * uu + dR_dr * srdc * uur * uu + duur_r);
            TIc[k] = (-1.0) * (de1_du1 * srxc * uur * uu + de2_du1 * srxc * uur * uu + dR_du1 * srdc * uur * uu +
                               e1n * e2m * srxc * uu * u2xrij[k] + e1n * e2m * srxc * uur * q.u2[k]);
            TJc[k] = (-1.0) * (de1_du2 * srxc * uur * uu + de2_du2 * srxc * uur * uu + dR_du2 * srdc * uur * uu +
                               e1n * e2m * srxc * uu * rijxu1[k] + e1n * e2m * srxc * uur * q.u1[k]);
        }
        double T1[3], T2[3], T1c[3], T2c[3];
        cross3(q.u1, TI, T1);
        cross3(q.u2, TJ, T2);
        cross3(q.u1, TIc, T1c);
        cross3(q.u2, TJc, T2c);
        pot = pot + chir * (e1n * e2m * srxc * uur * uu);
        for (int k = 0; k < 3; ++k) {
            o.f[k] = o.f[k] + chir * FC[k];
            o.ti[k] = T1[k] + chir * T1c[k];
            o.tj[k] = T2[k] + chir * T2c[k];
        }
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

// p = cutoff, mu, nu; chi' of particle a's energies is recomputed per pair as
// in the reference
template <int VARIANT>
__device__ __forceinline__ void d_gb(const DState &s, int a, int b, const double *p, PairOut &o)
{
    int mu = (int)p[1], nu = (int)p[2];
    const double wf = (VARIANT == 2 || VARIANT == 4) ? p[3] : 0.0;
    const double chir = VARIANT == 3 ? p[3] : (VARIANT == 4 ? p[4] : 0.0);
    double r[3], u2[3];
    GbSelf A;
    pair_rij(s, a, b, r);
    gb_self(s, a, mu, A);
    quat_u(s, b, u2);
    gb_core<VARIANT>(r, A, u2, p[0], mu, nu, o, wf, chir);
}

// The same pair with particle a's factors and particle b's axis read from a
// per-particle table of GbSelf records (GB_TAB doubles each: u, chi, chi',
// sig_x, eps_z), filled once per force evaluation by k_gb_self: the values
// d_gb computes per pair, without the two pow calls per pair when mu != 1.
#define GB_TAB 8
__device__ __forceinline__ void gb_self_store(const GbSelf &g, double *t)
{
    t[0] = g.u[0]; t[1] = g.u[1]; t[2] = g.u[2];
    t[3] = g.chi; t[4] = g.xhi; t[5] = g.lprp; t[6] = g.epsz; t[7] = 0.0;
}

template <int VARIANT>
__device__ __forceinline__ void d_gb_tab(const DState &s, int a, int b, const double *p, const double *tab,
                                         PairOut &o)
{
    int mu = (int)p[1], nu = (int)p[2];
    const double wf = (VARIANT == 2 || VARIANT == 4) ? p[3] : 0.0;
    const double chir = VARIANT == 3 ? p[3] : (VARIANT == 4 ? p[4] : 0.0);
    double r[3];
    pair_rij(s, a, b, r);
    const double *ta = tab + (long)GB_TAB * a, *tb = tab + (long)GB_TAB * b;
    GbSelf A;
    A.u[0] = ta[0]; A.u[1] = ta[1]; A.u[2] = ta[2];
    A.chi = ta[3]; A.xhi = ta[4]; A.lprp = ta[5]; A.epsz = ta[6];
    const double u2[3] = {tb[0], tb[1], tb[2]};
    gb_core<VARIANT>(r, A, u2, p[0], mu, nu, o, wf, chir);
}

// ---------------------------------------------------------------------------
// gb_ (variant 0) in the FP32 pair mode.  gb_ is ill-conditioned for FP32:
// for close pairs with aligned axes g(chi) cancels towards (1-chi)/(1+chi),
// sigma = sig_x / sqrt(g) carries that error into r - sigma + sigma_0, and the
// 12th power of sr amplifies it again, so rounding the inputs alone to FP32
// moves such a pair's force by ~3e-5 of its size; the torques' cross products
// with the axes cancel for the same pairs, and FP32 strength factors
// eps1^nu eps2^mu (a common factor of every term of a pair) leave ~2e-7 of
// each pair's force in particles whose pair forces cancel, up to 1.6e-5 of
// the particle's force in the batched workload.  So this form computes in
// FP64 throughout and saves what the FP32 mode is for, the long division and
// square-root sequences: g of chi and chi' with its gradients in r, u1 and u2
// (gb_g, gb_dgdr, gb_dgdu) share each division by r^2 and by 1 +- chi u1.u2,
// sigma and g^(-3/2) share one reciprocal square root, and eps1 is an FP32
// reciprocal square root refined by a Newton step: 8 FP64 divisions and 2
// square roots per pair (one more division in the switch band), where the
// reference-order form needs about 60 divisions.  The results are rounded to
// FP32 on output.
// ---------------------------------------------------------------------------
struct GbOri {
    double g, dgr[3], dgu1[3], dgu2[3];
};

__device__ __forceinline__ void gb_ori(double c, const GbGeo &q, double inv_rsq, GbOri &o)
{
    const double a = (q.ru1 + q.ru2) / (1.0 + c * q.uu), b = (q.ru1 - q.ru2) / (1.0 - c * q.uu);
    o.g = 1.0 - 0.5 * c * inv_rsq * (a * (q.ru1 + q.ru2) + b * (q.ru1 - q.ru2));
    const double h = 0.5 * c * c * inv_rsq * (a * a - b * b);
    for (int k = 0; k < 3; ++k) {
        o.dgr[k] = -c * inv_rsq * (a * (q.u1[k] + q.u2[k]) + b * (q.u1[k] - q.u2[k])) +
                   2.0 * q.r[k] * inv_rsq * (1.0 - o.g);
        o.dgu1[k] = -c * q.r[k] * inv_rsq * (a + b) + h * q.u2[k];
        o.dgu2[k] = -c * q.r[k] * inv_rsq * (a - b) + h * q.u1[k];
    }
}

__device__ __forceinline__ void gb_core_f32(const double r[3], const GbSelf &A, const double u2[3], double cutoff,
                                            int mu, int nu, PairOutT<float> &o)
{
    o.pe = 0.0f;
    for (int k = 0; k < 3; ++k) { o.f[k] = 0.0f; o.ti[k] = 0.0f; o.tj[k] = 0.0f; }
    GbGeo q;
    for (int k = 0; k < 3; ++k) { q.r[k] = r[k]; q.u1[k] = A.u[k]; q.u2[k] = u2[k]; }
    q.rsq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    q.ru1 = r[0] * q.u1[0] + r[1] * q.u1[1] + r[2] * q.u1[2];
    q.ru2 = r[0] * q.u2[0] + r[1] * q.u2[1] + r[2] * q.u2[2];
    q.uu = q.u1[0] * q.u2[0] + q.u1[1] * q.u2[1] + q.u1[2] * q.u2[2];
    const double rij_mag = sqrt(q.rsq), inv_mag = 1.0 / rij_mag, inv_rsq = inv_mag * inv_mag;
    GbOri C, X;
    gb_ori(A.chi, q, inv_rsq, C);
    const double rs = 1.0 / sqrt(C.g);
    const double sigma = A.lprp * rs;            // sigma_0 = 1
    const double rlim = sigma + cutoff - 1.0;
    if (!(rij_mag < rlim)) return;
    gb_ori(A.xhi, q, inv_rsq, X);
    const double sr = 1.0 / (rij_mag - sigma + 1.0);
    const double sr2 = sr * sr, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
    const double srx = sr12 - 2.0 * sr6, srd = 12.0 * (sr6 - sr12) * sr;
    const double g32 = sigma * rs * rs;          // sigma / g
    // strength factors: eps1 = (1 - (chi u1.u2)^2)^(-1/2) from the FP32
    // reciprocal square root and one Newton step in FP64, eps2 = g(chi') / eps_z
    const double x1 = 1.0 - (A.chi * q.uu) * (A.chi * q.uu);
    const double y0 = (double)__builtin_amdgcn_rsqf((float)x1);
    const double eps1 = y0 * (1.5 - 0.5 * x1 * y0 * y0);
    const double eps2 = X.g / A.epsz;
    const double e1n = d_ipow(eps1, nu), e2m = d_ipow(eps2, mu);
    const double e2m1 = d_ipow(eps2, mu - 1), e1n2 = d_ipow(eps1, nu + 2);
    const double E12 = e1n * e2m, DE2 = e1n * ((double)mu * e2m1);
    const double E1 = e2m * ((double)nu * e1n2) * (A.chi * A.chi * q.uu);
    double pot = E12 * srx, f[3], TI[3], TJ[3], ti[3], tj[3];
    for (int k = 0; k < 3; ++k) {
        const double dR_dr = r[k] * inv_mag + 0.5 * C.dgr[k] * g32;
        f[k] = -(E12 * srd * dR_dr + DE2 * X.dgr[k] * srx);
        TI[k] = -(srx * (E1 * q.u2[k] + DE2 * X.dgu1[k]) + E12 * (srd * 0.5 * C.dgu1[k] * g32));
        TJ[k] = -(srx * (E1 * q.u1[k] + DE2 * X.dgu2[k]) + E12 * (srd * 0.5 * C.dgu2[k] * g32));
    }
    cross3(q.u1, TI, ti);
    cross3(q.u2, TJ, tj);
    const double sw = 0.9 * rlim;
    if (rij_mag >= sw) {
        const double factor = 1.0 / (rlim - sw);
        const double xx = (rij_mag - sw) * factor;
        const double sx = 1.0 - 6.0 * xx * xx * xx * xx * xx + 15.0 * xx * xx * xx * xx - 10.0 * xx * xx * xx;
        const double sdx = -30.0 * xx * xx * xx * xx + 60.0 * xx * xx * xx - 30.0 * xx * xx;
        for (int k = 0; k < 3; ++k) {
            f[k] = sx * f[k] - sdx * pot * factor;
            ti[k] = sx * ti[k] - sdx * pot * factor;
            tj[k] = sx * tj[k] - sdx * pot * factor;
        }
        pot = sx * pot;
    }
    for (int k = 0; k < 3; ++k) { o.f[k] = (float)f[k]; o.ti[k] = (float)ti[k]; o.tj[k] = (float)tj[k]; }
    o.pe = (float)pot;
}

// d_gb_tab<0> in the FP32 pair mode
__device__ __forceinline__ void d_gb_tab_f32(const DState &s, int a, int b, const double *p, const double *tab,
                                             PairOutT<float> &o)
{
    double r[3];
    pair_rij(s, a, b, r);
    const double *ta = tab + (long)GB_TAB * a, *tb = tab + (long)GB_TAB * b;
    GbSelf A;
    A.u[0] = ta[0]; A.u[1] = ta[1]; A.u[2] = ta[2];
    A.chi = ta[3]; A.xhi = ta[4]; A.lprp = ta[5]; A.epsz = ta[6];
    const double u2[3] = {tb[0], tb[1], tb[2]};
    gb_core_f32(r, A, u2, p[0], (int)p[1], (int)p[2], o);
}

// ---------------------------------------------------------------------------
// point dipoles (dipole_calculate_forces, force.f90:2687-2751): the body z
// axis of each particle scaled by its charge; the reaction-field side effects
// of the reference routine are applied by a separate kernel
// ---------------------------------------------------------------------------
__device__ __forceinline__ void dip_axis(const DState &s, int i, double u[3])
{
    double w = s.f[SX_QW][i], x = s.f[SX_QX][i], y = s.f[SX_QY][i], z = s.f[SX_QZ][i];
    u[0] = 2.0 * (x * z + w * y);
    u[1] = 2.0 * (y * z - w * x);
    u[2] = w * w - x * x - y * y + z * z;
    double q = s.f[SX_CHARGE][i];
    u[0] = u[0] * q; u[1] = u[1] * q; u[2] = u[2] * q;
}

// true when the pair lies within the cutoff (the test of force.f90:2720)
__device__ __forceinline__ bool dip_in_range(const DState &s, int a, int b, double cutoff, double r[3],
                                             double &rij_sq, double &rij_mag)
{
    pair_rij(s, a, b, r);
    rij_sq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    rij_mag = sqrt(rij_sq);
    return rij_mag < cutoff;
}

__device__ __forceinline__ void d_dipole(const DState &s, int a, int b, const double *p, PairOut &o)
{
    double cutoff = p[0];
    double r[3], rij_sq, rij_mag;
    o.pe = 0.0;
    for (int k = 0; k < 3; ++k) o.f[k] = o.ti[k] = o.tj[k] = 0.0;
    if (!dip_in_range(s, a, b, cutoff, r, rij_sq, rij_mag)) return;
    double u1[3], u2[3];
    dip_axis(s, a, u1);
    dip_axis(s, b, u2);
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
    o.pe = (cij / rij_cube) - ((3.0 * ci * cj) / rij_pow5);
    double c1[3], c2[3];
    cross3(u1, t2, c1);
    cross3(u2, t3, c2);
    for (int k = 0; k < 3; ++k) {
        o.f[k] = (3.0 / rij_pow5) * t1[k];
        o.ti[k] = (1.0 / rij_cube) * c1[k];
        o.tj[k] = (1.0 / rij_cube) * c2[k];
    }
}

// isotropic dispatch
template <int KIND>
__device__ __forceinline__ void d_pair(const DState &s, int a, int b, const double *p, PairOut &o)
{
    if (KIND == SX_LJ) d_lj(s, a, b, p, o);
    else if (KIND == SX_LJ12) d_lj12(s, a, b, p, o);
    else if (KIND == SX_LJW) d_ljw(s, a, b, p, o);
    else if (KIND == SX_WCA) d_wca(s, a, b, p, o);
    else if (KIND == SX_MORSE) d_morse(s, a, b, p, o);
    else if (KIND == SX_DH) d_dh(s, a, b, p, o);
    else if (KIND == SX_GB) d_gb<0>(s, a, b, p, o);
    else if (KIND == SX_GBR) d_gb<1>(s, a, b, p, o);
    else if (KIND == SX_GBW) d_gb<2>(s, a, b, p, o);
    else if (KIND == SX_GBC) d_gb<3>(s, a, b, p, o);
    else if (KIND == SX_GBWC) d_gb<4>(s, a, b, p, o);
    else if (KIND == SX_DIPOLE) d_dipole(s, a, b, p, o);
    else if (KIND == SX_COULOMB) d_coulomb(s, a, b, p, o);
}

// the isotropic kinds of the FP32 pair mode
template <int KIND>
constexpr bool sx_has_f32()
{
    return KIND == SX_LJ || KIND == SX_WCA || KIND == SX_DH;
}

// d_pair in the arithmetic type T (float: the kinds of sx_has_f32 only)
template <int KIND, typename T>
__device__ __forceinline__ void d_pair_t(const DState &s, int a, int b, const double *p, PairOutT<T> &o)
{
    if constexpr (sizeof(T) == sizeof(double)) {
        d_pair<KIND>(s, a, b, p, o);
    } else {
        static_assert(sx_has_f32<KIND>(), "no FP32 form of this pair kind");
        double r[3];
        pair_rij(s, a, b, r);
        if constexpr (KIND == SX_LJ) lj_core<float>(r, p, o);
        else if constexpr (KIND == SX_WCA) wca_core<float>(r, p, o);
        else dh_core<float>(r, s.f[SX_CHARGE][a], s.f[SX_CHARGE][b], p, o);
    }
}
