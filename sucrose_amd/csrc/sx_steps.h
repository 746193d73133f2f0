// Per-particle building blocks of the integrators (integrators.f90), shared by
// the single-system kernels and the batched step.
#pragma once
#include "sx_pairs.h"

__device__ __forceinline__ void wrap_pos(const DState &s, int i)
{
    if (s.periodic) {
        double x = s.f[SX_RX][i], y = s.f[SX_RY][i], z = s.f[SX_RZ][i];
        s.f[SX_RX][i] = sx_mimg(x, s.box[0]);
        s.f[SX_RY][i] = sx_mimg(y, s.box[1]);
        s.f[SX_RZ][i] = sx_mimg(z, s.box[2]);
    }
}

__device__ __forceinline__ void half_kick(const DState &s, int i, double dt)
{
    double im = s.f[SX_INV_MASS][i];
    s.f[SX_VX][i] = s.f[SX_VX][i] + 0.5 * dt * (s.f[SX_FX][i] * im);
    s.f[SX_VY][i] = s.f[SX_VY][i] + 0.5 * dt * (s.f[SX_FY][i] * im);
    s.f[SX_VZ][i] = s.f[SX_VZ][i] + 0.5 * dt * (s.f[SX_FZ][i] * im);
}

__device__ __forceinline__ void save_old(const DState &s, int i)
{
    s.f[SX_RX_OLD][i] = s.f[SX_RX][i];
    s.f[SX_RY_OLD][i] = s.f[SX_RY][i];
    s.f[SX_RZ_OLD][i] = s.f[SX_RZ][i];
}

__device__ __forceinline__ double kin_trans_of(double mass, double vx, double vy, double vz)
{
    return 0.5 * mass * (vx * vx + vy * vy + vz * vz);
}

__device__ __forceinline__ double kin_trans(const DState &s, int i)
{
    return kin_trans_of(s.f[SX_MASS][i], s.f[SX_VX][i], s.f[SX_VY][i], s.f[SX_VZ][i]);
}

// RM = R(q)^T (integrators.f90:144-152)
__device__ __forceinline__ void rm_of(double w, double x, double y, double z, double RM[9])
{
    RM[0] = w * w + x * x - y * y - z * z;   // (1,1)
    RM[3] = 2.0 * (x * y + w * z);           // (1,2)
    RM[6] = 2.0 * (x * z - w * y);           // (1,3)
    RM[1] = 2.0 * (x * y - w * z);           // (2,1)
    RM[4] = w * w - x * x + y * y - z * z;   // (2,2)
    RM[7] = 2.0 * (y * z + w * x);           // (2,3)
    RM[2] = 2.0 * (x * z + w * y);           // (3,1)
    RM[5] = 2.0 * (y * z - w * x);           // (3,2)
    RM[8] = w * w - x * x - y * y + z * z;   // (3,3)
}

// q = (w, x, y, z), l the angular momentum, ii the inverse principal moments
__device__ __forceinline__ double kin_rot_of(const double q[4], double lx, double ly, double lz, const double ii[3])
{
    double RM[9];
    rm_of(q[0], q[1], q[2], q[3], RM);
    double bx = (RM[0] * lx + RM[3] * ly + RM[6] * lz);
    double by = (RM[1] * lx + RM[4] * ly + RM[7] * lz);
    double bz = (RM[2] * lx + RM[5] * ly + RM[8] * lz);
    return 0.5 * ((bx * bx) * ii[0] + (by * by) * ii[1] + (bz * bz) * ii[2]);
}

__device__ __forceinline__ double kin_rot(const DState &s, int i)
{
    const double q[4] = {s.f[SX_QW][i], s.f[SX_QX][i], s.f[SX_QY][i], s.f[SX_QZ][i]};
    const double ii[3] = {s.f[SX_IXX_INV][i], s.f[SX_IYY_INV][i], s.f[SX_IZZ_INV][i]};
    return kin_rot_of(q, s.f[SX_LX][i], s.f[SX_LY][i], s.f[SX_LZ][i], ii);
}

__device__ __forceinline__ void quat_midpoint(const DState &s, int i, double dt, double &qdw, double &qdx, double &qdy,
                                              double &qdz, double RM[9], double &bx, double &by, double &bz,
                                              bool inertia_after)
{
    double qw0 = s.f[SX_QW][i], qx0 = s.f[SX_QX][i], qy0 = s.f[SX_QY][i], qz0 = s.f[SX_QZ][i];
    double lx = s.f[SX_LX][i], ly = s.f[SX_LY][i], lz = s.f[SX_LZ][i];
    double iix = s.f[SX_IXX_INV][i], iiy = s.f[SX_IYY_INV][i], iiz = s.f[SX_IZZ_INV][i];
    double qwn = qw0, qxn = qx0, qyn = qy0, qzn = qz0;
    for (int pass = 0; pass < 2; ++pass) {
        double qwo = qwn, qxo = qxn, qyo = qyn, qzo = qzn;
        rm_of(qwo, qxo, qyo, qzo, RM);
        double wx, wy, wz;
        bx = (RM[0] * lx + RM[3] * ly + RM[6] * lz);
        by = (RM[1] * lx + RM[4] * ly + RM[7] * lz);
        bz = (RM[2] * lx + RM[5] * ly + RM[8] * lz);
        // qq_ multiplies the bracket by I^-1 in one expression; lgv_ stores L_body first:
        // the same two roundings either way
        wx = bx * iix; wy = by * iiy; wz = bz * iiz;
        (void)inertia_after;
        qdw = 0.5 * (-qxo * wx - qyo * wy - qzo * wz);
        qdx = 0.5 * (qwo * wx + qyo * wz - qzo * wy);
        qdy = 0.5 * (qwo * wy - qxo * wz + qzo * wx);
        qdz = 0.5 * (qwo * wz + qxo * wy - qyo * wx);
        qwn = qw0 + 0.5 * dt * qdw;
        qxn = qx0 + 0.5 * dt * qdx;
        qyn = qy0 + 0.5 * dt * qdy;
        qzn = qz0 + 0.5 * dt * qdz;
    }
}

__device__ __forceinline__ void quat_advance(const DState &s, int i, double dt, double qdw, double qdx, double qdy,
                                             double qdz)
{
    double w = s.f[SX_QW][i] + dt * qdw, x = s.f[SX_QX][i] + dt * qdx;
    double y = s.f[SX_QY][i] + dt * qdy, z = s.f[SX_QZ][i] + dt * qdz;
    double qm = sqrt(w * w + x * x + y * y + z * z);
    s.f[SX_QW][i] = w / qm; s.f[SX_QX][i] = x / qm; s.f[SX_QY][i] = y / qm; s.f[SX_QZ][i] = z / qm;
}

__device__ __forceinline__ void ang_kick(const DState &s, int i, double dt)
{
    s.f[SX_LX][i] = s.f[SX_LX][i] + 0.5 * dt * s.f[SX_TX][i];
    s.f[SX_LY][i] = s.f[SX_LY][i] + 0.5 * dt * s.f[SX_TY][i];
    s.f[SX_LZ][i] = s.f[SX_LZ][i] + 0.5 * dt * s.f[SX_TZ][i];
}

struct LgvConst {
    double dt, damp, T, e1, e2, kk1, kk2;   // e1 = exp(-damp dt), e2 = 1 - exp(-2 damp dt)
    int temp_int, save;
};

// lgv_initial_step (integrators.f90:507-627) for particle i; ctr is the
// particle's 1-based index in its system (the first Philox counter word)
__device__ __forceinline__ void lgv_initial_particle(const DState &s, int i, int ctr, const LgvConst &k, int st)
{
    double dt = k.dt;
    half_kick(s, i, dt);
    if (k.save) save_old(s, i);
    double im = s.f[SX_INV_MASS][i];
    double rx = s.f[SX_RX][i] + 0.5 * dt * s.f[SX_VX][i];
    double ry = s.f[SX_RY][i] + 0.5 * dt * s.f[SX_VY][i];
    double rz = s.f[SX_RZ][i] + 0.5 * dt * s.f[SX_VZ][i];
    double n1x, n1y, n1z, n2x, n2y, n2z;
    sx_normal3(ctr, st, k.temp_int, 0, n1x, n1y, n1z);
    sx_normal3(ctr, st, k.temp_int, 1, n2x, n2y, n2z);
    double amp = sqrt(k.T * im * k.e2);
    double vx = k.e1 * s.f[SX_VX][i] + amp * n1x;
    double vy = k.e1 * s.f[SX_VY][i] + amp * n1y;
    double vz = k.e1 * s.f[SX_VZ][i] + amp * n1z;
    s.f[SX_VX][i] = vx; s.f[SX_VY][i] = vy; s.f[SX_VZ][i] = vz;
    s.f[SX_RX][i] = rx + 0.5 * dt * vx;
    s.f[SX_RY][i] = ry + 0.5 * dt * vy;
    s.f[SX_RZ][i] = rz + 0.5 * dt * vz;
    wrap_pos(s, i);
    ang_kick(s, i, dt);
    double qdw, qdx, qdy, qdz, RM[9], bx, by, bz;
    quat_midpoint(s, i, dt, qdw, qdx, qdy, qdz, RM, bx, by, bz, false);
    bx = k.kk2 * k.kk1 * bx + k.kk2 * sqrt(2.0 * k.damp * k.T * dt * s.f[SX_IXX][i]) * n2x;
    by = k.kk2 * k.kk1 * by + k.kk2 * sqrt(2.0 * k.damp * k.T * dt * s.f[SX_IYY][i]) * n2y;
    bz = k.kk2 * k.kk1 * bz + k.kk2 * sqrt(2.0 * k.damp * k.T * dt * s.f[SX_IZZ][i]) * n2z;
    s.f[SX_LX][i] = (RM[0] * bx + RM[1] * by + RM[2] * bz);
    s.f[SX_LY][i] = (RM[3] * bx + RM[4] * by + RM[5] * bz);
    s.f[SX_LZ][i] = (RM[6] * bx + RM[7] * by + RM[8] * bz);
    quat_advance(s, i, dt, qdw, qdx, qdy, qdz);
}

