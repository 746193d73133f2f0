// Composite rigid bodies (system.f90:743-881): copy_transform places a
// satellite at its centre plus a body-frame offset rotated by the centre's
// orientation and gives it the centre's quaternion; transfer_forces moves a
// satellite's force and torque, plus the torque of that force about the
// centre, onto the centre.
//
// On the device one thread handles one placement, or one centre with all of
// its satellites in the order the caller gave them, so every centre receives
// its satellites' contributions in the reference's sequence.  The inertia
// bookkeeping of copy_transform(..., set_inertia = .true.) is a set-up step
// that runs on the host in call order.
#include "sx_internal.h"

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

__global__ void k_copy_transform(DState s, int count, const int *from, const int *to, const double *off)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    int a = from[t], b = to[t];
    double w = s.f[SX_QW][a], x = s.f[SX_QX][a], y = s.f[SX_QY][a], z = s.f[SX_QZ][a];
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
    double xo = off[3L * t], yo = off[3L * t + 1], zo = off[3L * t + 2];
    s.f[SX_RX][b] = s.f[SX_RX][a] + (rot[0][0] * xo + rot[0][1] * yo + rot[0][2] * zo);
    s.f[SX_RY][b] = s.f[SX_RY][a] + (rot[1][0] * xo + rot[1][1] * yo + rot[1][2] * zo);
    s.f[SX_RZ][b] = s.f[SX_RZ][a] + (rot[2][0] * xo + rot[2][1] * yo + rot[2][2] * zo);
    s.f[SX_QW][b] = w;
    s.f[SX_QX][b] = x;
    s.f[SX_QY][b] = y;
    s.f[SX_QZ][b] = z;
}

// centre c = centres[k] receives satellites sats[start[k] .. start[k+1])
__global__ void k_transfer_forces(DState s, int ncentres, const int *centres, const int *start, const int *sats)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncentres) return;
    int to = centres[k];
    double F[3] = {s.f[SX_FX][to], s.f[SX_FY][to], s.f[SX_FZ][to]};
    double T[3] = {s.f[SX_TX][to], s.f[SX_TY][to], s.f[SX_TZ][to]};
    const int fi[3] = {SX_FX, SX_FY, SX_FZ}, ti[3] = {SX_TX, SX_TY, SX_TZ}, ri[3] = {SX_RX, SX_RY, SX_RZ};
    for (int q = start[k]; q < start[k + 1]; ++q) {
        int from = sats[q];
        double r[3], f[3], t[3];
        for (int d = 0; d < 3; ++d) {
            r[d] = s.f[ri[d]][from] - s.f[ri[d]][to];
            if (s.periodic) r[d] = sx_mimg(r[d], s.box[d]);
            f[d] = s.f[fi[d]][from];
        }
        t[0] = r[1] * f[2] - r[2] * f[1];
        t[1] = r[2] * f[0] - r[0] * f[2];
        t[2] = r[0] * f[1] - r[1] * f[0];
        for (int d = 0; d < 3; ++d) {
            F[d] = F[d] + f[d];
            T[d] = T[d] + s.f[ti[d]][from] + t[d];
            s.f[fi[d]][from] = 0.0;
            s.f[ti[d]][from] = 0.0;
        }
    }
    for (int d = 0; d < 3; ++d) {
        s.f[fi[d]][to] = F[d];
        s.f[ti[d]][to] = T[d];
    }
}

namespace {

// checks a call's index arrays (1-based in, 0-based out): the written atoms
// distinct and disjoint from the read ones, so that the parallel update
// equals the reference's sequence of calls
void check_pairs(sx_ctx *c, int count, const int *a1, const int *b1, std::vector<int> &a, std::vector<int> &b,
                 const char *written)
{
    if (count < 0 || (count > 0 && (!a1 || !b1))) throw sx_error{SX_ERR_ARG, "bad composite arrays"};
    a.resize(count);
    b.resize(count);
    std::vector<char> role(c->n, 0);   // 1 read, 2 written
    for (int t = 0; t < count; ++t) {
        a[t] = a1[t] - 1;
        b[t] = b1[t] - 1;
        if (a[t] < 0 || a[t] >= c->n || b[t] < 0 || b[t] >= c->n || a[t] == b[t])
            throw sx_error{SX_ERR_ARG, "Error: Atom index does not exist in system"};
    }
    const std::vector<int> &w = std::string(written) == "to" ? b : a;
    const std::vector<int> &r = std::string(written) == "to" ? a : b;
    for (int t = 0; t < count; ++t) {
        if (role[w[t]] == 2) throw sx_error{SX_ERR_ARG, "an atom is written twice in one call"};
        role[w[t]] = 2;
    }
    for (int t = 0; t < count; ++t)
        if (role[r[t]] == 2) throw sx_error{SX_ERR_ARG, "an atom is both read and written in one call"};
}

// the inertia bookkeeping of copy_transform (system.f90:773-840), in order
void set_inertia_host(sx_ctx *c, const std::vector<int> &from, const std::vector<int> &to, const double *off)
{
    int n = c->n;
    const int fl[] = {SX_MASS, SX_SIG_X, SX_SIG_Y, SX_SIG_Z, SX_IXX, SX_IYY, SX_IZZ};
    std::vector<double> h[7];
    for (int k = 0; k < 7; ++k) {
        h[k].resize(n);
        HIPCHK(hipMemcpyAsync(h[k].data(), c->d_fields + (long)fl[k] * c->ld, sizeof(double) * n,
                              hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    auto &mass = h[0], &sx = h[1], &sy = h[2], &sz = h[3], &ixx = h[4], &iyy = h[5], &izz = h[6];
    auto from_shape = [&](int i) {
        ixx[i] = (std::pow(sy[i], 2.0) + std::pow(sz[i], 2.0)) / 20.0;
        iyy[i] = (std::pow(sx[i], 2.0) + std::pow(sz[i], 2.0)) / 20.0;
        izz[i] = (std::pow(sx[i], 2.0) + std::pow(sy[i], 2.0)) / 20.0;
    };
    std::string err;
    for (size_t t = 0; t < from.size() && err.empty(); ++t) {
        int a = from[t], b = to[t];
        double xo = off[3 * t], yo = off[3 * t + 1], zo = off[3 * t + 2];
        if (!(mass[b] <= 0.0)) {
            err = "Error: Second particle " + std::to_string(b + 1) + "must have mass zero";
            break;
        }
        bool b_has_i = ixx[b] > 0.0 && iyy[b] > 0.0 && izz[b] > 0.0;
        bool b_has_shape = sx[b] > 0.0 && sy[b] > 0.0 && sz[b] > 0.0;
        if (ixx[a] > 0.0 && iyy[a] > 0.0 && izz[a] > 0.0) {
        } else if (sx[a] > 0.0 && sy[a] > 0.0 && sz[a] > 0.0) {
            from_shape(a);
        } else {
            err = "Error: First particle " + std::to_string(a + 1) + "must have non-zero shape";
            break;
        }
        if (b_has_i) {
        } else if (b_has_shape) {
            from_shape(b);
        } else {
            err = "Error: Second particle " + std::to_string(b + 1) + "must have non-zero shape";
            break;
        }
        ixx[a] = ixx[a] + ixx[b] + (std::pow(yo, 2.0) + std::pow(zo, 2.0));
        iyy[a] = iyy[a] + iyy[b] + (std::pow(xo, 2.0) + std::pow(zo, 2.0));
        izz[a] = izz[a] + izz[b] + (std::pow(xo, 2.0) + std::pow(yo, 2.0));
    }
    for (int k = 4; k < 7; ++k)
        HIPCHK(hipMemcpyAsync(c->d_fields + (long)fl[k] * c->ld, h[k].data(), sizeof(double) * n,
                              hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!err.empty()) throw sx_error{SX_ERR_ARG, err};
}

template <class T>
T *upload(sx_ctx *c, const std::vector<T> &v)
{
    T *d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(T) * std::max<size_t>(1, v.size())));
    if (!v.empty()) HIPCHK(hipMemcpy(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return d;
}

// centre-major tables of a set of (satellite -> centre) transfers, keeping
// each centre's satellites in call order
void transfer_tables(const std::vector<int> &sat, const std::vector<int> &centre, std::vector<int> &centres,
                     std::vector<int> &start, std::vector<int> &sats)
{
    std::vector<int> order(sat.size());
    for (size_t t = 0; t < sat.size(); ++t) order[t] = (int)t;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return centre[x] < centre[y]; });
    centres.clear();
    start.assign(1, 0);
    sats.clear();
    for (int t : order) {
        if (centres.empty() || centres.back() != centre[t]) {
            if (!centres.empty()) start.push_back((int)sats.size());
            centres.push_back(centre[t]);
        }
        sats.push_back(sat[t]);
    }
    start.push_back((int)sats.size());
    if (centres.empty()) start.assign(1, 0);
}

}  // namespace

void sx_composite_free(sx_ctx *c)
{
    Composites &m = c->comp;
    void *p[] = {m.centre, m.sat, m.off, m.t_centres, m.t_start, m.t_sats};
    for (void *q : p)
        if (q) hipFree(q);
    m = Composites();
}

void sx_composite_place_impl(sx_ctx *c)
{
    Composites &m = c->comp;
    if (m.count == 0) return;
    hipLaunchKernelGGL(k_copy_transform, dim3(sx_blocks(m.count, 256)), dim3(256), 0, c->stream, c->dstate(),
                       m.count, m.centre, m.sat, m.off);
}

void sx_composite_transfer_impl(sx_ctx *c)
{
    Composites &m = c->comp;
    if (m.count == 0) return;
    hipLaunchKernelGGL(k_transfer_forces, dim3(sx_blocks(m.ncentres, 256)), dim3(256), 0, c->stream, c->dstate(),
                       m.ncentres, m.t_centres, m.t_start, m.t_sats);
}

extern "C" {

int sx_copy_transform(sx_ctx *c, int count, const int *from1, const int *to1, const double *offsets, int set_inertia)
{
    return guard(c, [&] {
        need_system(c);
        std::vector<int> a, b;
        check_pairs(c, count, from1, to1, a, b, "to");
        if (count == 0) return;
        if (!offsets) throw sx_error{SX_ERR_ARG, "null offsets"};
        std::vector<double> off(offsets, offsets + 3L * count);
        int *da = upload(c, a), *db = upload(c, b);
        double *doff = upload(c, off);
        hipLaunchKernelGGL(k_copy_transform, dim3(sx_blocks(count, 256)), dim3(256), 0, c->stream, c->dstate(), count,
                           da, db, doff);
        HIPCHK(hipStreamSynchronize(c->stream));
        hipFree(da); hipFree(db); hipFree(doff);
        if (set_inertia == 1) set_inertia_host(c, a, b, offsets);
    });
}

int sx_transfer_forces(sx_ctx *c, int count, const int *from1, const int *to1)
{
    return guard(c, [&] {
        need_system(c);
        std::vector<int> a, b;
        check_pairs(c, count, from1, to1, a, b, "from");
        if (count == 0) return;
        std::vector<int> centres, start, sats;
        transfer_tables(a, b, centres, start, sats);
        int *dc = upload(c, centres), *ds = upload(c, start), *dt = upload(c, sats);
        hipLaunchKernelGGL(k_transfer_forces, dim3(sx_blocks((int)centres.size(), 256)), dim3(256), 0, c->stream,
                           c->dstate(), (int)centres.size(), dc, ds, dt);
        HIPCHK(hipStreamSynchronize(c->stream));
        hipFree(dc); hipFree(ds); hipFree(dt);
    });
}

int sx_set_composites(sx_ctx *c, int count, const int *centre1, const int *sat1, const double *offsets)
{
    return guard(c, [&] {
        need_system(c);
        std::vector<int> a, b;
        check_pairs(c, count, centre1, sat1, a, b, "to");
        if (count > 0 && !offsets) throw sx_error{SX_ERR_ARG, "null offsets"};
        std::vector<int> centres, start, sats;
        transfer_tables(b, a, centres, start, sats);
        sx_composite_free(c);
        Composites &m = c->comp;
        m.count = count;
        m.ncentres = (int)centres.size();
        m.centre = upload(c, a);
        m.sat = upload(c, b);
        m.off = upload(c, std::vector<double>(offsets, offsets + 3L * count));
        m.t_centres = upload(c, centres);
        m.t_start = upload(c, start);
        m.t_sats = upload(c, sats);
        c->have_graph = false;
    });
}

int sx_composite_place(sx_ctx *c)
{
    return guard(c, [&] { need_system(c); sx_composite_place_impl(c); });
}

int sx_composite_transfer(sx_ctx *c)
{
    return guard(c, [&] { need_system(c); sx_composite_transfer_impl(c); });
}

}  // extern "C"
