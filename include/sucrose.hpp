// sucrose.hpp -- C++ face over the C ABI of sucrose.h, shaped like the
// reference's C++ classes (include/force.h:8-73, include/integrators.h:8-36)
// so that code written against `Force::calculate_forces(i, j)` and
// `Integrator::initial_step()/final_step()` can drive the device instead.
//
// Differences from the reference's C++ face, all in the direction of the
// reference's Fortran modules (the C++ face is an older port: its
// LennardJones has no cutoff, its LJ12 and WCA take the running virial as a
// force coefficient, its Verlet step ignores the mass):
//   * potentials take the cutoff the Fortran routines take (the C++
//     LennardJones has none);
//   * indices are 0-based here, as in the C++ face, and converted to the
//     C ABI's 1-based convention;
//   * calculate_forces(i, j) evaluates one pair on the device and adds it to
//     the device state (F(i) += f, F(j) -= f, pot_eng += pe, torques for the
//     anisotropic kinds) with a few small copies per pair: a porting and
//     checking path; calculate_all() replaces the whole pair loop and is the
//     fast path.
// Every failing call throws sucrose::Error with the library's message.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "sucrose.h"

namespace sucrose {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(sx_ctx *ctx, int rc)
{
    if (rc != SX_OK) throw Error(rc, sx_last_error(ctx) ? sx_last_error(ctx) : "sucrose error");
}

// one device-resident system (the reference's System / module system)
class Device {
  public:
    explicit Device(int device = 0)
    {
        int rc = sx_create(device, &ctx_);
        if (rc != SX_OK) throw Error(rc, sx_last_error(nullptr));
    }
    ~Device() { if (ctx_) sx_destroy(ctx_); }
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;

    sx_ctx *ctx() const { return ctx_; }
    int size() const { return n_; }

    void set_num_atoms(int n, int rigid = 1) { check(ctx_, sx_set_num_atoms(ctx_, n, rigid)); n_ = n * rigid; }
    void set_box(double lx, double ly, double lz, bool periodic = true)
    {
        double b[3] = {lx, ly, lz};
        check(ctx_, sx_set_box(ctx_, b, periodic ? 1 : 0));
    }
    void upload(int field, const std::vector<double> &v) { check(ctx_, sx_upload(ctx_, field, v.data(), (int)v.size())); }
    std::vector<double> download(int field) const
    {
        std::vector<double> v(n_);
        check(ctx_, sx_download(ctx_, field, v.data(), n_));
        return v;
    }
    double scalar(int id) const
    {
        double v = 0;
        check(ctx_, sx_get_scalar(ctx_, id, &v));
        return v;
    }
    void set_scalar(int id, double v) { check(ctx_, sx_set_scalar(ctx_, id, v)); }
    void initialize_system(bool remove_com = true) { check(ctx_, sx_initialize_system(ctx_, remove_com ? 1 : 0)); }
    void init_nlist(double skin, double cutoff) { check(ctx_, sx_nlist_init(ctx_, skin, cutoff)); }
    void update_nlist() { check(ctx_, sx_nlist_update(ctx_)); }
    void initialize_forces() { check(ctx_, sx_initialize_forces(ctx_)); }
    void calculate_thermo() { check(ctx_, sx_calculate_thermo(ctx_)); }

  private:
    sx_ctx *ctx_ = nullptr;
    int n_ = 0;
};

// Force (include/force.h:7-17)
class Force {
  protected:
    Device &system;
    int kind;
    std::vector<double> params;
    Force(Device &sys, int k, std::vector<double> p) : system(sys), kind(k), params(std::move(p)) {}

  public:
    virtual ~Force() = default;
    // one pair, 0-based, added to the device state
    virtual void calculate_forces(int i, int j)
    {
        int i1 = i + 1, j1 = j + 1;
        double pe = 0, f[3], ti[3], tj[3], lam = 0;
        sx_ctx *c = system.ctx();
        check(c, sx_pair_eval(c, kind, params.data(), (int)params.size(), 1, &i1, &j1, &pe, f, ti, tj, &lam));
        add_pair(i, j, pe, f, ti, tj);
    }
    // the whole pair loop over the neighbour list (README.md:86-103)
    void calculate_all() { check(system.ctx(), sx_forces(system.ctx(), kind, params.data(), (int)params.size())); }

  private:
    void add_pair(int i, int j, double pe, const double f[3], const double ti[3], const double tj[3])
    {
        sx_ctx *c = system.ctx();
        const int ff[3] = {SX_FX, SX_FY, SX_FZ}, tf[3] = {SX_TX, SX_TY, SX_TZ};
        for (int k = 0; k < 3; ++k) {
            bump(c, ff[k], i, f[k]);
            bump(c, ff[k], j, -f[k]);
            bump(c, tf[k], i, ti[k]);
            bump(c, tf[k], j, tj[k]);
        }
        double p = 0;
        check(c, sx_get_scalar(c, SX_POT_ENG, &p));
        check(c, sx_set_scalar(c, SX_POT_ENG, p + pe));
    }
    // one element, read, incremented and written back through the C ABI
    static void bump(sx_ctx *c, int field, int i, double d)
    {
        if (d == 0.0) return;
        double cur = 0.0;
        check(c, sx_get_atom_values(c, field, i + 1, 1, &cur));
        cur = cur + d;
        check(c, sx_set_atom_values(c, field, i + 1, 1, &cur));
    }
};

class LennardJones : public Force {       // include/force.h:19-30; lj_calculate_forces
  public:
    LennardJones(Device &sys, double epsilon, double sigma, double cutoff)
        : Force(sys, SX_LJ, {sigma, epsilon, cutoff}) {}
};
class LennardJones12 : public Force {     // :32-42; lj12_calculate_forces
  public:
    LennardJones12(Device &sys, double epsilon, double sigma, double cutoff)
        : Force(sys, SX_LJ12, {sigma, epsilon, cutoff}) {}
};
class LennardJones_WCA : public Force {   // :44-53; wca_calculate_forces
  public:
    LennardJones_WCA(Device &sys, double epsilon, double sigma) : Force(sys, SX_WCA, {sigma, epsilon}) {}
};
class MorsePotential : public Force {     // :55-67; morse_calculate_forces
  public:
    MorsePotential(Device &sys, double dissociation_energy, double equilibrium_distance, double width, double cutoff)
        : Force(sys, SX_MORSE, {dissociation_energy, width, equilibrium_distance, cutoff}) {}
};
class GayBernePotential : public Force {  // :69-73 (empty in the C++ face); gb_calculate_forces
  public:
    GayBernePotential(Device &sys, double cutoff, int mu = 1, int nu = 2)
        : Force(sys, SX_GB, {cutoff, (double)mu, (double)nu}) {}
};

// Integrator (include/integrators.h:8-24)
class Integrator {
  protected:
    Device &system;
    double dt;

  public:
    Integrator(Device &sys, double dt_) : system(sys), dt(dt_) {}
    virtual ~Integrator() = default;
    double get_timestep() const { return dt; }
    void set_timestep(double d) { dt = d; }
    virtual void initial_step() = 0;
    virtual void final_step() = 0;
};

class VerletIntegrator : public Integrator {   // include/integrators.h:26-34; vv_initial/final_step
  public:
    VerletIntegrator(Device &sys, double dt_) : Integrator(sys, dt_) {}
    void initial_step() override { check(system.ctx(), sx_vv_initial_step(system.ctx(), dt)); }
    void final_step() override { check(system.ctx(), sx_vv_final_step(system.ctx(), dt)); }
};

class LangevinIntegrator : public Integrator {  // lgv_initial/final_step (integrators.f90:507, :629)
  public:
    LangevinIntegrator(Device &sys, double dt_, double damping) : Integrator(sys, dt_), gamma(damping) {}
    void initial_step() override { check(system.ctx(), sx_lgv_initial_step(system.ctx(), dt, ++step, gamma)); }
    void final_step() override { check(system.ctx(), sx_lgv_final_step(system.ctx(), dt, gamma)); }

  private:
    double gamma;
    int step = 0;
};

}  // namespace sucrose
