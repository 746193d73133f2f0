// Whole-step driver: one step of the README loop (README.md:83-115) captured
// once as a HIP graph and replayed.  The Langevin step counter and the
// neighbour-list rebuild decision live on the device, so replays need no host
// round trip.
#include "sx_internal.h"

#include <cstdlib>
#include <cstring>

__global__ void k_step_set(int *dstep, int v) { *dstep = v; }

struct DeferPe {
    sx_ctx *c;
    explicit DeferPe(sx_ctx *c_) : c(c_) { c->defer_pe = true; }
    ~DeferPe() { c->defer_pe = false; c->ecp_pp_by_step = false; }
};

static bool is_ecp(int kind)
{
    return kind == SX_ECP || kind == SX_ECPR || kind == SX_ECPW || kind == SX_ECPC || kind == SX_ECPWC;
}

static void enqueue_step(sx_ctx *c, const sx_step_spec &sp)
{
    DeferPe defer(c);   // pot_eng is reduced by the final-step kernel
    if (sp.integrator == SX_INT_LGV) {
        // initial step, rebuild decision, initialize_forces (and the ECP
        // per-particle records) in one kernel, then the conditional rebuild
        bool ecp = false;
        for (int f = 0; f < sp.nforces; ++f) ecp = ecp || is_ecp(sp.kinds[f]);
        sx_lgv_step_begin_impl(c, sp.dt, sp.damping, ecp);
        c->ecp_pp_by_step = ecp;
        sx_nlist_build_impl(c, true);
    } else {
        switch (sp.integrator) {
        case SX_INT_VV: sx_vv_initial_impl(c, sp.dt, false); break;
        case SX_INT_NHT:
            sx_nht_impl(c, sp.dt, nullptr, 0, false);
            sx_qq_initial_impl(c, sp.dt);
            break;
        case SX_INT_LGVC: sx_lgv_initial_impl(c, sp.dt, 0, c->d_step, sp.damping, true); break;
        }
        sx_nlist_flag_impl(c);
        sx_nlist_build_impl(c, true);
        sx_init_forces_impl(c);
    }
    for (int f = 0; f < sp.nforces; ++f) {
        int np = 0;
        switch (sp.kinds[f]) {
        case SX_LJ: case SX_LJ12: case SX_DH: np = 3; break;
        case SX_LJW: case SX_MORSE: np = 4; break;
        case SX_WCA: np = 2; break;
        case SX_GB: case SX_GBR: np = 3; break;
        case SX_GBW: case SX_GBC: np = 4; break;
        case SX_GBWC: np = 5; break;
        case SX_ECP: np = 1; break;
        case SX_ECPW: case SX_ECPC: np = 2; break;
        case SX_ECPWC: np = 3; break;
        default: np = 0;
        }
        sx_forces_impl(c, sp.kinds[f], sp.params[f], np);
    }
    if (sp.bonded) sx_bonded_impl(c);
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
        if (!c->d_step) HIPCHK(hipMalloc(&c->d_step, sizeof(int)));
        hipLaunchKernelGGL(k_step_set, dim3(1), dim3(1), 0, c->stream, c->d_step, spec->first_step);
        if (nsteps == 0) return;
        // SUCROSE_NO_GRAPH=1 enqueues the same launches eagerly (per-kernel tracing tools)
        const char *ng = getenv("SUCROSE_NO_GRAPH");
        if (ng && ng[0] == '1') {
            for (int t = 0; t < nsteps; ++t) enqueue_step(c, *spec);
            sx_check_device_error(c);
            return;
        }
        sx_step_spec key = *spec;
        key.first_step = 0;
        if (!c->have_graph || !same_spec(key, c->graph_spec)) {
            HIPCHK(hipStreamSynchronize(c->stream));
            if (c->graph_exec) { hipGraphExecDestroy(c->graph_exec); c->graph_exec = nullptr; }
            if (c->graph) { hipGraphDestroy(c->graph); c->graph = nullptr; }
            // one eager step sizes every lazily allocated buffer
            sx_step_spec warm = *spec;
            enqueue_step(c, warm);
            sx_check_device_error(c);
            hipLaunchKernelGGL(k_step_set, dim3(1), dim3(1), 0, c->stream, c->d_step, spec->first_step + 1);
            --nsteps;
            HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            try {
                enqueue_step(c, *spec);
            } catch (...) {
                hipGraph_t g;
                hipStreamEndCapture(c->stream, &g);
                if (g) hipGraphDestroy(g);
                throw;
            }
            HIPCHK(hipStreamEndCapture(c->stream, &c->graph));
            HIPCHK(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
            c->graph_spec = key;
            c->have_graph = true;
        }
        for (int t = 0; t < nsteps; ++t) HIPCHK(hipGraphLaunch(c->graph_exec, c->stream));
        sx_check_device_error(c);
    });
}
