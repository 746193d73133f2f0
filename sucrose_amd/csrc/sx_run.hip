// Whole-step driver: one step of the README loop (README.md:83-115) captured
// once as a HIP graph and replayed.  The Langevin step counter and the
// neighbour-list rebuild decision live on the device, so replays need no host
// round trip.
#include "sx_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

__global__ void k_step_set(int *dstep, int v) { *dstep = v; }

struct DeferPe {
    sx_ctx *c;
    explicit DeferPe(sx_ctx *c_) : c(c_) { c->defer_pe = true; }
    ~DeferPe()
    {
        c->defer_pe = false;
        c->ecp_pp_by_step = false;
        c->ecp_apply_in_final = false;
    }
};

// the parameter count of each pair kind in a step spec
static int nparams_of(int kind)
{
    switch (kind) {
    case SX_LJ: case SX_LJ12: case SX_DH: return 3;
    case SX_LJW: case SX_MORSE: return 4;
    case SX_WCA: return 2;
    case SX_GB: case SX_GBR: return 3;
    case SX_GBW: case SX_GBC: return 4;
    case SX_GBWC: return 5;
    case SX_ECP: return 1;
    case SX_ECPW: case SX_ECPC: return 2;
    case SX_ECPWC: return 3;
    case SX_DIPOLE: case SX_COULOMB: return 2;
    default: return 0;
    }
}

static bool is_ecp(int kind)
{
    return kind == SX_ECP || kind == SX_ECPR || kind == SX_ECPW || kind == SX_ECPC || kind == SX_ECPWC;
}

static void enqueue_step(sx_ctx *c, const sx_step_spec &sp)
{
    DeferPe defer(c);   // pot_eng is reduced by the final-step kernel
    if (sp.integrator == SX_INT_LGV && c->comp.count == 0) {
        // initial step, rebuild decision, initialize_forces (and the ECP
        // per-particle records) in one kernel, then the conditional rebuild
        bool ecp = false;
        for (int f = 0; f < sp.nforces; ++f) ecp = ecp || is_ecp(sp.kinds[f]);
        sx_lgv_step_begin_impl(c, sp.dt, sp.damping, ecp);
        c->ecp_pp_by_step = ecp;
        sx_nlist_build_impl(c, true);
    } else {
        // without composite bodies the Nose-Hoover step's quaternion half,
        // the rebuild decision and initialize_forces are one kernel
        const bool fused = sp.integrator == SX_INT_NHT && c->comp.count == 0;
        switch (sp.integrator) {
        case SX_INT_VV: sx_vv_initial_impl(c, sp.dt, false); break;
        case SX_INT_NHT:
            sx_nht_impl(c, sp.dt, nullptr, 0, false);
            if (fused) sx_qq_step_begin_impl(c, sp.dt);
            else sx_qq_initial_impl(c, sp.dt);
            break;
        case SX_INT_LGV: sx_lgv_initial_impl(c, sp.dt, 0, c->d_step, sp.damping, false); break;
        case SX_INT_LGVC: sx_lgv_initial_impl(c, sp.dt, 0, c->d_step, sp.damping, true); break;
        }
        if (!fused) {
            sx_composite_place_impl(c);   // satellites follow their centres
            sx_nlist_flag_impl(c);
        }
        sx_nlist_build_impl(c, true);
        if (!fused) sx_init_forces_impl(c);
    }
    bool dipoles = false;
    for (int f = 0; f < sp.nforces; ++f) dipoles = dipoles || sp.kinds[f] == SX_DIPOLE;
    const bool two = sp.nforces == 2 &&
                     sx_forces_two_impl(c, sp.kinds[0], sp.params[0], nparams_of(sp.kinds[0]), sp.kinds[1],
                                        sp.params[1], nparams_of(sp.kinds[1]));
    // an ECP term that adds the last forces of the step (no bonded terms,
    // reaction field or composite transfers after it) is applied by the
    // Langevin final-step kernel itself
    const bool apply_in_final = !two && sp.integrator == SX_INT_LGV && c->comp.count == 0 && sp.nforces > 0 &&
                                is_ecp(sp.kinds[sp.nforces - 1]) && !sp.bonded && !dipoles;
    for (int f = 0; f < sp.nforces && !two; ++f) {
        c->ecp_apply_in_final = apply_in_final && f == sp.nforces - 1;
        sx_forces_impl(c, sp.kinds[f], sp.params[f], nparams_of(sp.kinds[f]));
    }
    if (dipoles) sx_eval_reaction_field_impl(c);
    if (sp.bonded) sx_bonded_impl(c);
    sx_composite_transfer_impl(c);   // satellite forces onto their centres
    // final step with calculate_thermo folded in (and the Langevin counter advance)
    switch (sp.integrator) {
    case SX_INT_VV: sx_vv_final_impl(c, sp.dt, true); break;
    case SX_INT_NHT:
        sx_nht_impl(c, sp.dt, nullptr, 1, false);
        sx_qq_final_impl(c, sp.dt, true);
        break;
    case SX_INT_LGV: sx_lgv_final_impl(c, sp.dt, sp.damping, false, true, c->d_step); break;
    case SX_INT_LGVC: sx_lgv_final_impl(c, sp.dt, sp.damping, true, true, c->d_step); break;
    }
}

static bool same_spec(const sx_step_spec &a, const sx_step_spec &b)
{
    return memcmp(&a, &b, sizeof(sx_step_spec)) == 0;
}

// steps per captured graph: a replay launch costs ~13 us on its own and each
// kernel node ~1 us more, so a run replays the longest chunks that fit
static const int SX_GRAPH_CHUNK[4] = {1, 4, 16, 64};

static void drop_graphs(sx_ctx *c)
{
    for (int k = 0; k < 4; ++k) {
        if (c->graph_exec[k]) { hipGraphExecDestroy(c->graph_exec[k]); c->graph_exec[k] = nullptr; }
        if (c->graph[k]) { hipGraphDestroy(c->graph[k]); c->graph[k] = nullptr; }
    }
    c->have_graph = false;
}

// captures SX_GRAPH_CHUNK[k] consecutive steps as one graph
static void capture_chunk(sx_ctx *c, const sx_step_spec &sp, int k)
{
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    try {
        for (int t = 0; t < SX_GRAPH_CHUNK[k]; ++t) enqueue_step(c, sp);
    } catch (...) {
        hipGraph_t g = nullptr;
        hipStreamEndCapture(c->stream, &g);
        if (g) hipGraphDestroy(g);
        throw;
    }
    HIPCHK(hipStreamEndCapture(c->stream, &c->graph[k]));
    HIPCHK(hipGraphInstantiate(&c->graph_exec[k], c->graph[k], nullptr, nullptr, 0));
}

extern "C" int sx_run(sx_ctx *c, const sx_step_spec *spec, int nsteps)
{
    return guard(c, [&] {
        need_system(c);
        if (!spec || nsteps < 0 || spec->nforces < 0 || spec->nforces > 4) throw sx_error{SX_ERR_ARG, "bad step spec"};
        if (!c->nl.init) throw sx_error{SX_ERR_STATE, "neighbour list not initialised (sx_nlist_init)"};
        if (spec->integrator != SX_INT_VV && !c->have_target_temp)
            throw sx_error{SX_ERR_NOTEMP, "Error: Target temperature not set."};
        for (int f = 0; f < spec->nforces; ++f)
            if (is_ecp(spec->kinds[f])) sx_nlist_want_half(c);
        // first_step < 0: continue from the device's Langevin counter (the
        // step after the last one run, or the one a checkpoint restored)
        const bool resume = spec->first_step < 0;
        if (resume && !c->d_step)
            throw sx_error{SX_ERR_STATE, "no Langevin step counter to continue from (first_step < 0)"};
        if (!c->d_step) HIPCHK(hipMalloc(&c->d_step, sizeof(int)));
        if (!resume) hipLaunchKernelGGL(k_step_set, dim3(1), dim3(1), 0, c->stream, c->d_step, spec->first_step);
        if (nsteps == 0) return;
        // SUCROSE_NO_GRAPH=1 enqueues the same launches eagerly (per-kernel tracing tools)
        const char *ng = getenv("SUCROSE_NO_GRAPH");
        if (ng && ng[0] == '1') {
            sx_nlist_headroom(c);
            for (int t = 0; t < nsteps; ++t) enqueue_step(c, *spec);
            sx_nlist_check_after_replay(c, true);
            return;
        }
        sx_step_spec key = *spec;
        key.first_step = 0;
        if (!c->have_graph || !same_spec(key, c->graph_spec)) {
            HIPCHK(hipStreamSynchronize(c->stream));
            drop_graphs(c);
            // room for the list to grow inside the replays, then one eager
            // step sizes every lazily allocated buffer
            sx_nlist_headroom(c);
            sx_step_spec warm = *spec;
            enqueue_step(c, warm);
            sx_nlist_check_after_replay(c, false);
            if (!resume)
                hipLaunchKernelGGL(k_step_set, dim3(1), dim3(1), 0, c->stream, c->d_step, spec->first_step + 1);
            --nsteps;
            c->graph_spec = key;
            c->have_graph = true;
        }
        // the remaining steps as the longest captured chunks that fit
        for (int k = 3; k >= 0; --k) {
            while (nsteps >= SX_GRAPH_CHUNK[k]) {
                if (!c->graph_exec[k]) capture_chunk(c, *spec, k);
                HIPCHK(hipGraphLaunch(c->graph_exec[k], c->stream));
                nsteps -= SX_GRAPH_CHUNK[k];
            }
        }
        sx_nlist_check_after_replay(c, true);
    });
}

// ---------------------------------------------------------------------------
// Monte Carlo barostat (integrators.f90:864-945)
// ---------------------------------------------------------------------------
__global__ void k_scale_positions(DState s, double scale, int save)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    if (save) {   // RX_old = RX (integrators.f90:898-900)
        s.f[SX_RX_OLD][i] = s.f[SX_RX][i];
        s.f[SX_RY_OLD][i] = s.f[SX_RY][i];
        s.f[SX_RZ_OLD][i] = s.f[SX_RZ][i];
    } else {      // RX = RX_old * scale (:917-919; scale 1 restores, :937-939)
        s.f[SX_RX][i] = s.f[SX_RX_OLD][i] * scale;
        s.f[SX_RY][i] = s.f[SX_RY_OLD][i] * scale;
        s.f[SX_RZ][i] = s.f[SX_RZ_OLD][i] * scale;
    }
}

// compute_potential: the list rebuilt at the current box, initialize_forces,
// the pair loops (and bonded terms) of the spec; returns pot_eng
static double mc_potential(sx_ctx *c, const sx_step_spec &sp)
{
    sx_nlist_rebox_impl(c);
    sx_init_forces_impl(c);
    for (int f = 0; f < sp.nforces; ++f) sx_forces_impl(c, sp.kinds[f], sp.params[f], nparams_of(sp.kinds[f]));
    if (sp.bonded) sx_bonded_impl(c);
    sx_sync_scalars_to_host(c);
    return c->h_scal[SX_POT_ENG];
}

extern "C" int sx_set_target_pressure(sx_ctx *c, double pres)
{
    return guard(c, [&] { c->target_pres = pres; });
}

extern "C" int sx_mc_barostat(sx_ctx *c, const sx_step_spec *spec, double volume_change, const double u[2],
                              int *accept)
{
    return guard(c, [&] {
        need_system(c);
        if (!spec || !u || spec->nforces < 0 || spec->nforces > 4) throw sx_error{SX_ERR_ARG, "bad arguments"};
        if (!c->nl.init) throw sx_error{SX_ERR_STATE, "neighbour list not initialised (sx_nlist_init)"};
        if (c->target_pres <= 1e-3) throw sx_error{SX_ERR_ARG, "Error: Target pressure too small."};
        if (accept) *accept = 0;
        double max_dV = volume_change < 0.0 ? 0.01 : volume_change;
        sx_sync_scalars_to_host(c);
        double beta = 1.0 / c->h_scal[SX_TEMP];
        double V_old = c->box[0] * c->box[1] * c->box[2];
        double L_old[3] = {c->box[0], c->box[1], c->box[2]};
        DState s = c->dstate();
        int nb = sx_blocks(c->n, 256);
        hipLaunchKernelGGL(k_scale_positions, dim3(nb), dim3(256), 0, c->stream, s, 1.0, 1);
        // forces and torques of the state the move starts from, restored on
        // a rejected move (with the list and pot_eng of that state)
        long n = c->n;
        std::vector<double> keep(6 * n);
        const int fk[6] = {SX_FX, SX_FY, SX_FZ, SX_TX, SX_TY, SX_TZ};
        for (int k = 0; k < 6; ++k)
            HIPCHK(hipMemcpyAsync(keep.data() + k * n, c->d_fields + (long)fk[k] * c->ld, sizeof(double) * n,
                                  hipMemcpyDeviceToHost, c->stream));
        double scal_keep[SX_NSCALARS];
        std::copy(c->h_scal, c->h_scal + SX_NSCALARS, scal_keep);
        auto restore = [&]() {
            for (int d = 0; d < 3; ++d) c->box[d] = L_old[d];
            DState s0 = c->dstate();
            hipLaunchKernelGGL(k_scale_positions, dim3(nb), dim3(256), 0, c->stream, s0, 1.0, 0);
            sx_nlist_rebox_impl(c);
            for (int k = 0; k < 6; ++k)
                HIPCHK(hipMemcpyAsync(c->d_fields + (long)fk[k] * c->ld, keep.data() + k * n, sizeof(double) * n,
                                      hipMemcpyHostToDevice, c->stream));
            std::copy(scal_keep, scal_keep + SX_NSCALARS, c->h_scal);
            sx_sync_scalars_to_device(c);
            HIPCHK(hipStreamSynchronize(c->stream));
        };
        double U_old = mc_potential(c, *spec);
        double V_new = V_old + max_dV * (2.0 * u[0] - 1.0) * V_old;
        if (V_new <= 0.0) {   // unphysical: no move (integrators.f90:907-909)
            restore();
            return;
        }
        double scale = std::pow(V_new / V_old, 1.0 / 3.0);
        for (int d = 0; d < 3; ++d) c->box[d] = L_old[d] * scale;
        hipLaunchKernelGGL(k_scale_positions, dim3(nb), dim3(256), 0, c->stream, c->dstate(), scale, 0);
        double U_new;
        try {
            U_new = mc_potential(c, *spec);
        } catch (const sx_error &e) {
            restore();
            throw;
        }
        double dU = U_new - U_old;
        double exponent = -beta * dU - beta * c->target_pres * (V_new - V_old) + (double)c->n * std::log(V_new / V_old);
        if (std::log(u[1]) < exponent) {
            double volume = c->box[0] * c->box[1] * c->box[2];
            c->h_scal[SX_DENSITY] = (double)c->n / volume;
            sx_sync_scalars_to_device(c);
            HIPCHK(hipStreamSynchronize(c->stream));
            if (accept) *accept = 1;
        } else {
            restore();
        }
        c->have_graph = false;
    });
}
