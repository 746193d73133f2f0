#!/usr/bin/env python3
"""Benchmark: SUCROSE ECP Langevin step on one MI355X.

Workload (north-star configuration): 32768 prolate ellipsoids (semi-axes
0.5, 0.5, 1.5), packing fraction 0.35, repulsive ECP (ecpr_calculate_forces,
force.f90:1387) with the reference's Brent contact solve, translational and
rotational Langevin (lgv_initial_step / lgv_final_step, integrators.f90:507,
:629) at T* = 1, gamma = 1, dt = 0.002, Philox4x32-10 noise, neighbour list
cutoff 3.1 + skin 0.3.  One "step" = initial step, list update, force
evaluation, final step, thermo -- the reference driver loop (README.md:83-115).
Inputs are resident in HBM; the timed region replays a captured HIP graph.

Prints one JSON line (metric/value/unit/... plus roofline and cpu_baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s on 1 MI355X (GB, ECP); learned vs plain ECP at 1e-5 force err"
PEAK_FP64_TFLOPS = 78.6
# algorithmic FP64 operation counts (add, mul, div, sqrt, round = 1 each)
FLOP_OPTIM_DIST = 62        # one optim_dist evaluation (blend + 3x3 Cholesky solve + dot)
FLOP_BRENT_ITER = 20        # Brent bookkeeping per iteration
FLOP_SCREEN = 30 + FLOP_OPTIM_DIST   # min image, bound test, f(psi)
FLOP_ECP_FORCE = 650        # force/torque evaluation at the solved lambda
FLOP_APPLY_ENTRY = 7        # per solved pair and particle in the per-particle apply
# L2-miss bytes per launch of each ECP stage kernel at this workload, from
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (KiB per dispatch) scaled by a 256 MiB
# copy measured the same way (FETCH x2.000, WRITE x1.000); the runs and the
# arithmetic are in profiles/ecp_langevin_32768_fetch_write.txt
TRAFFIC_BYTES = {
    "ecp_precompute": (1694.6 * 2.0 + 8192.0) * 1024,
    "ecp_screen": (9848.3 * 2.0 + 866.6) * 1024,
    "ecp_solve": (10435.6 * 2.0 + 7626.9) * 1024,
    "ecp_apply": (10120.0 * 2.0 + 2816.0) * 1024,
}
# FP64 operations issued per launch of each ECP stage kernel at this workload:
# rocprofv3 --pmc SQ_INSTS_VALU_FLOPS_FP64 (counted per wavefront instruction,
# FMA = 2, transcendental included) times the 64 lanes of a wavefront; the run
# is in profiles/ecp_langevin_32768_fp64_counters.txt (the screen computes in
# FP32 and issues no FP64 operations)
ISSUED_FP64 = {
    "ecp_precompute": 1.224e5 * 64,
    "ecp_screen": 0.0,
    "ecp_solve": 2.565e6 * 64,
    "ecp_apply": 3.515e4 * 64,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """The host CPU's model name and the CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        ncpu = os.cpu_count() or 1
    return model, ncpu


def cpu_all_cores(state_arrays, box, params, seconds):
    """The C restatement of the same loop with its pair terms and list build
    spread over all host threads (liboracle_omp.so; results equal the serial
    loop's bit for bit): the all-cores CPU figure, kind "port"."""
    from oracle.oracle import ECPR, FIELDS, LGV, Oracle, State
    path = os.path.join(ROOT, "oracle", "liboracle_omp.so")
    if not os.path.exists(path):
        return None
    n = params["n"]
    st = State(n, [box, box, box])
    for k, f in enumerate(FIELDS):
        st.buf[:, k] = state_arrays[f.upper()]
    st["target_temp"] = params["T"]
    st["dof"] = 6 * n - 3
    st["density"] = n / box ** 3
    T = Oracle(path)
    nl = T.nlist_new(st, params["cutoff"], params["skin"])
    T.nlist_build_threads(st, nl)
    dt, gamma, step0 = params["dt"], params["gamma"], params["next_step"]
    steps = 0
    t0 = time.time()
    while time.time() - t0 < seconds or steps == 0:
        T.integrate(LGV, 0, st, dt, step=step0 + steps, damping=gamma)
        if T.needs_rebuild(st, nl):
            T.nlist_build_threads(st, nl)
        T.call("or_initialize_forces", st)
        T.pair_loop_threads(st, nl, ECPR, [0.0])
        T.integrate(LGV, 1, st, dt, damping=gamma)
        T.call("or_calculate_thermo", st)
        steps += 1
    el = time.time() - t0
    return {"value": n * steps / el, "unit": "particle-steps/s", "cores": T.threads(), "kind": "port",
            "sample": f"{steps} steps of the C restatement of the reference loop (lgv + ecpr pair loop over its "
                      f"Verlet list, reference rebuild rule) on the same {n}-particle state, pair terms and list "
                      f"build over {T.threads()} OpenMP threads, accumulation in the loop's order",
            "seconds": el}


def chains_cpu_baseline(state, p, seconds):
    """The reference's own charged-chain loop (oracle/ref_harness.f90
    ref_run_chain_steps: lgvc_initial_step with RATTLE, its Verlet list, WCA
    and Debye-Hueckel over it, cosine_bend_ and dih_ terms, lgvc_final_step
    with RATTLE, calculate_thermo) on the device's state, in a private copy of
    the reference library."""
    import numpy as np
    from oracle.oracle import FIELDS, Reference, State, reference_available
    if not reference_available(fast=True):
        return None
    n, L = p["n"], p["box"]
    st = State(n, [L, L, L])
    for k, f in enumerate(FIELDS):
        st.buf[:, k] = state[f.upper()]
    ids = np.arange(n).reshape(p["nchains"], p["nbeads"])
    st.bbi = ids[:, :-1].ravel().astype(np.int32)
    st.bbj = ids[:, 1:].ravel().astype(np.int32)
    st.bb_len = np.full(len(st.bbi), 1.0)
    st["target_temp"] = p["T"]
    st["dof"] = 6 * n - st.nc - 3
    st["density"] = n / L ** 3
    bends = np.stack([ids[:, :-2], ids[:, 1:-1], ids[:, 2:]], -1).reshape(-1, 3)
    dihs = np.stack([ids[:, :-3], ids[:, 1:-2], ids[:, 2:-1], ids[:, 3:]], -1).reshape(-1, 4)
    wca = p["forces"][0][1:3]
    dh = p["forces"][1][1:4]
    R = Reference(fast=True, private=True)
    R.load(st)
    t0 = time.time()
    R.nlist_init(p["skin"], dh[0])
    log(f"[chains cpu] reference init_nlist {time.time() - t0:.1f}s")
    step = p["next_step"]
    t0 = time.time()
    kb, ab, kd = p["k_bend"], p["bend_angle"], p["k_dih"]
    R.run_chain_steps(1, step, p["dt"], p["gamma"], wca, dh, bends, kb, ab, dihs, kd)
    t1 = max(time.time() - t0, 1e-3)
    k = int(max(1, min(30, seconds / t1)))
    t0 = time.time()
    R.run_chain_steps(k, step + 1, p["dt"], p["gamma"], wca, dh, bends, kb, ab, dihs, kd)
    el = time.time() - t0
    return {"value": n * k / el, "unit": "particle-steps/s", "cores": 1, "kind": "reference",
            "sample": f"{k} steps of the reference's charged-chain loop (lgvc + RATTLE, wca and debye_huckel over "
                      f"its Verlet list, cosine_bend_ and dih_ terms) on the same {n}-bead state after one untimed "
                      f"step", "seconds": el}


def cpu_baseline(state_arrays, box, params, seconds):
    """The reference's own loop (built from source, oracle/_ref) on the host;
    falls back to the C restatement when the reference build is absent."""
    import numpy as np
    from oracle.oracle import ECPR, FIELDS, LGV, Oracle, Reference, State, reference_available
    n = params["n"]
    st = State(n, [box, box, box])
    for k, f in enumerate(FIELDS):
        st.buf[:, k] = state_arrays[f.upper()]
    st["target_temp"] = params["T"]
    st["dof"] = 6 * n - 3
    st["density"] = n / box ** 3
    dt, gamma, step0 = params["dt"], params["gamma"], params["next_step"]
    if reference_available(fast=True):
        R = Reference(fast=True)
        R.load(st)
        t0 = time.time()
        R.nlist_init(params["skin"], params["cutoff"])
        log(f"[cpu] reference init_nlist {time.time() - t0:.1f}s")
        t0 = time.time()
        R.run_steps(LGV, ECPR, [0.0], 1, step0, dt, gamma)
        t1 = time.time() - t0
        more = int(max(1, min(60, seconds / max(t1, 1e-3) - 1)))
        t0 = time.time()
        R.run_steps(LGV, ECPR, [0.0], more, step0 + 1, dt, gamma)
        el = time.time() - t0
        steps = more
        kind = "reference"
        sample = (f"{steps} steps of the reference loop (lgv + ecpr pair loop over its Verlet list, "
                  f"reference rebuild rule) on the same {n}-particle state after one untimed step; "
                  f"reference fortran built from source with flang -O3 -ffast-math -march=x86-64-v3")
    else:
        O = Oracle()
        nl = O.nlist_new(st, params["cutoff"], params["skin"])
        O.nlist_build(st, nl)
        t0 = time.time()
        steps = 0
        while time.time() - t0 < seconds or steps == 0:
            O.integrate(LGV, 0, st, dt, step=step0 + steps, damping=gamma)
            if O.needs_rebuild(st, nl):
                O.nlist_build(st, nl)
            O.call("or_initialize_forces", st)
            O.pair_loop(st, nl, ECPR, [0.0])
            O.integrate(LGV, 1, st, dt, damping=gamma)
            O.call("or_calculate_thermo", st)
            steps += 1
        el = time.time() - t0
        kind = "port"
        sample = f"{steps} steps of the C restatement of the reference loop on the same {n}-particle state"
    return {"value": n * steps / el, "unit": "particle-steps/s", "cores": 1, "kind": kind, "sample": sample,
            "seconds": el}


def gb_cpu_baseline(state, box, params, nsteps):
    """The reference's own NHT + Gay-Berne loop on the GB state, in a private
    copy of the reference library (the reference hard-codes GB(mu=1, nu=2);
    the cost per pair is the same)."""
    from oracle.oracle import FIELDS, GB, NHT, Reference, State, reference_available
    if not reference_available(fast=True):
        return None
    n = params["n"]
    st = State(n, [box, box, box])
    for k, f in enumerate(FIELDS):
        st.buf[:, k] = state[f.upper()]
    st["target_temp"] = params["T"]
    st["dof"] = 6 * n - 3
    st["density"] = n / box ** 3
    R = Reference(fast=True, private=True)
    R.load(st)
    R.nlist_init(params["skin"], params["cutoff"])
    R.run_steps(NHT, GB, [params["cutoff"]], 1, 1, params["dt"], 1.0)
    t0 = time.time()
    R.run_steps(NHT, GB, [params["cutoff"]], nsteps, 2, params["dt"], 1.0)
    el = time.time() - t0
    return {"value": n * nsteps / el, "unit": "particle-steps/s", "cores": 1, "kind": "reference",
            "sample": f"{nsteps} steps of the reference loop (nht + qq steps, gb pair loop over its Verlet list) on "
                      f"the same {n}-particle state after one untimed step", "seconds": el}


def elem_err(a, b):
    """Worst |a - b| / (|b| + median of the non-zero |b|) over the elements."""
    import numpy as np
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nz = np.abs(b[b != 0])
    med = float(np.median(nz)) if nz.size else 1.0
    return float(np.max(np.abs(a - b) / (np.abs(b) + med)))


FORCE_FIELDS = ("FX", "FY", "FZ", "TX", "TY", "TZ")


def f32_error(d, snap, evaluate, fields=FORCE_FIELDS, groups=None):
    """The FP32 pair terms' error on one state: evaluate() at 64 and at 32
    bits from the same snapshot (the 64-bit terms are the reference's
    arithmetic); worst force/torque element error (within each group of
    particles, e.g. each batched system, when groups are given) and relative
    energy error.  Leaves the snapshot loaded and the precision at 32 bits."""
    import numpy as np
    got = {}
    for bits in (64, 32):
        d.load_arrays(snap)
        d.set_pair_precision(bits)
        pe = evaluate()
        got[bits] = (np.stack([d.download(f) for f in fields]), pe)
    d.load_arrays(snap)
    f64, pe64 = got[64]
    f32, pe32 = got[32]
    groups = groups or [slice(None)]
    per = [max([elem_err(f32[k][g], f64[k][g]) for k in range(len(fields)) if np.any(f64[k][g])] or [0.0])
           for g in groups]
    worst = max(per)
    return {"force_elem_err": worst, "per_group": per,
            "pe_rel_err": float(np.max(np.abs(np.asarray(pe32) - np.asarray(pe64)) /
                                       np.maximum(1.0, np.abs(np.asarray(pe64))))),
            "measure": "FP32 vs FP64 pair terms on one state; |a-b|/(|b|+median|b|) per force/torque element"
                       + (", per system" if len(groups) > 1 else "")}


def f32_row(rec, n, steps, el, err, rebuilds, extra=None):
    """The FP32 companion of a workload's FP64 record."""
    row = {"value": n * steps / el, "unit": "particle-steps/s", "steps": steps, "ms_per_step": el / steps * 1e3,
           "speedup_vs_f64": rec["ms_per_step"] / (el / steps * 1e3), "rebuilds": rebuilds, "error": err,
           "config": dict(rec["config"], dtype="f32 pair terms (f64 positions, sums, integrator)")}
    if extra:
        row.update(extra)
    return row


def other_workloads(sx, args):
    """The Gay-Berne + Nose-Hoover, the charged-chain and the batched-replica
    configurations, each prepared on the device and timed over whole
    graph-replayed steps (a section with 0 steps is skipped); each is timed
    again with FP32 pair terms from the same state."""
    from sucrose_amd.workloads import chains_state, gb_nhc_state, load_prepared, save_prepared

    def prepared(d, name, make):
        # the workload from --load-states, else prepared (and kept with --save-states)
        if args.load_states:
            p = load_prepared(d, os.path.join(args.load_states, name + ".npz"))
            log(f"[workload] loaded {name} from {args.load_states}")
            return p
        p = make(d, log=log)
        if args.save_states:
            os.makedirs(args.save_states, exist_ok=True)
            save_prepared(d, p, os.path.join(args.save_states, name + ".npz"))
        return p

    out = {}
    if args.gb_steps > 0:
        d = sx.System(0)
        p = prepared(d, "gb_nhc_8000", gb_nhc_state)
        gb = [(sx.GB, p["cutoff"], p["mu"], p["nu"])]
        step = p["next_step"]
        d.run(sx.INT_NHT, gb, 100, p["dt"], first_step=step)
        d.synchronize()
        step += 100
        # the timed segment's starting point, so that the FP32 run can retrace it
        start = d.state()
        start_scal = {k: d.get_scalar(k) for k in sx.SCALARS}
        rb0 = d.rebuild_count()
        it0, calls0, fb0 = d.nht_stats()
        t0 = time.perf_counter()
        d.run(sx.INT_NHT, gb, args.gb_steps, p["dt"], first_step=step)
        d.synchronize()
        el = time.perf_counter() - t0
        it1, calls1, fb1 = d.nht_stats()
        rec = {"value": p["n"] * args.gb_steps / el, "unit": "particle-steps/s", "steps": args.gb_steps,
               "ms_per_step": el / args.gb_steps * 1e3, "rebuilds": d.rebuild_count() - rb0,
               "T": d.get_scalar("TEMP"), "pe_per_particle": d.get_scalar("POT_ENG") / p["n"],
               "chain_newton_updates_per_call": (it1 - it0) / max(1, calls1 - calls0),
               "chain_serial_fallbacks": fb1 - fb0,
               "config": {"workload": "gb_nhc_8000", "n": p["n"], "rho": p["rho"], "gb": "GB(3,5,2,1)",
                          "sig": [1, 1, 3], "eps": [5, 5, 1], "T": p["T"], "dt": p["dt"], "cutoff": p["cutoff"],
                          "skin": p["skin"], "list_radius": p["cutoff"] + p["skin"],
                          "integrator": "nht (chain 3, tau 2) + quaternion", "dtype": "f64"}}
        log(f"[gb] {args.gb_steps} steps in {el:.3f}s -> {rec['value']:.4e} particle-steps/s; "
            f"{rec['chain_newton_updates_per_call']:.2f} Newton updates per chain call, {fb1 - fb0} serial fallbacks")
        snap = d.state()
        if not args.no_cpu:
            cpu = gb_cpu_baseline(snap, p["box"], p, 10)
            if cpu:
                rec["cpu_baseline"] = cpu
                log(f"[gb cpu] reference {cpu['value']:.4e} particle-steps/s ({cpu['seconds']:.1f}s)")
        out["gb_nhc_8000"] = rec

        if not args.no_variants:   # the FP32 rows (and the list-radius variant)
            def gb_forces():
                d.initialize_forces()
                d.forces(*gb[0])
                return d.get_scalar("POT_ENG")

            err = f32_error(d, snap, gb_forces)
            err.pop("per_group")

            def back_to_start():
                d.load_arrays(start)
                for k, v in start_scal.items():
                    d.set_scalar(k, v)
                d.build_nlist()
                d.synchronize()

            def segment(bits, list_cutoff):
                """The timed segment again from its starting state, with the pair
                terms at `bits` and a list of radius list_cutoff + skin, after a
                warm-up run from the same state that captures the step graphs."""
                d.set_pair_precision(bits)
                d.init_nlist(p["skin"], list_cutoff)
                back_to_start()
                d.run(sx.INT_NHT, gb, 100, p["dt"], first_step=step)
                back_to_start()
                rb0 = d.rebuild_count()
                it0, calls0, fb0 = d.nht_stats()
                t0 = time.perf_counter()
                d.run(sx.INT_NHT, gb, args.gb_steps, p["dt"], first_step=step)
                d.synchronize()
                el = time.perf_counter() - t0
                it1, calls1, fb1 = d.nht_stats()
                return el, {"rebuilds": d.rebuild_count() - rb0, "T": d.get_scalar("TEMP"),
                            "chain_newton_updates_per_call": (it1 - it0) / max(1, calls1 - calls0),
                            "chain_serial_fallbacks": fb1 - fb0}

            el, x = segment(32, p["cutoff"])
            out["gb_nhc_8000_f32"] = f32_row(rec, p["n"], args.gb_steps, el, err, x.pop("rebuilds"), x)
            log(f"[gb f32] {el / args.gb_steps * 1e3:.4f} ms/step; error {err['force_elem_err']:.2e}; "
                f"{x['chain_newton_updates_per_call']:.2f} Newton updates per chain call, "
                f"{x['chain_serial_fallbacks']} serial fallbacks")
            # the list radius that covers gb_'s whole range, sigma_max + cutoff -
            # sigma_0 = 6.0, plus the skin (the headline rows use cutoff + skin)
            full = 6.0
            for bits in (64, 32):
                el, x = segment(bits, full)
                key = f"gb_nhc_8000_list{full + p['skin']:.1f}" + ("_f32" if bits == 32 else "")
                out[key] = dict(value=p["n"] * args.gb_steps / el, unit="particle-steps/s", steps=args.gb_steps,
                                ms_per_step=el / args.gb_steps * 1e3, **x,
                                config=dict(rec["config"], list_radius=full + p["skin"],
                                            dtype="f64" if bits == 64 else
                                            "f32 pair terms (f64 positions, sums, integrator)"))
                log(f"[gb list {full + p['skin']:.1f} {bits}-bit] {el / args.gb_steps * 1e3:.4f} ms/step")
        d.close()
    if args.chain_steps > 0:
        d = sx.System(0)
        p = prepared(d, "chains_100000", chains_state)
        warm = 64 + 16 + 4   # captures every step-graph length a long run replays
        step = p["next_step"]
        d.run(sx.INT_LGVC, p["forces"], warm, p["dt"], damping=p["gamma"], first_step=step, bonded=True)
        step += warm
        d.synchronize()
        start = d.state()
        start_scal = {k: d.get_scalar(k) for k in sx.SCALARS}
        rb0 = d.rebuild_count()
        t0 = time.perf_counter()
        d.run(sx.INT_LGVC, p["forces"], args.chain_steps, p["dt"], damping=p["gamma"], first_step=step,
              bonded=True)
        d.synchronize()
        el = time.perf_counter() - t0
        rec = {"value": p["n"] * args.chain_steps / el, "unit": "particle-steps/s", "steps": args.chain_steps,
               "ms_per_step": el / args.chain_steps * 1e3, "rebuilds": d.rebuild_count() - rb0,
               "T": d.get_scalar("TEMP"), "pe_per_particle": d.get_scalar("POT_ENG") / p["n"],
               "config": {"workload": "chains_100000", "chains": p["nchains"], "beads": p["nbeads"], "rho": p["rho"],
                          "T": p["T"], "dt": p["dt"], "forces": "wca(1,1) + debye_huckel(rc 5, kappa 1, eps_r 1)",
                          "bonded": "cosine_bend k=20 about straight + dih k(1+cos) k=1", "constraints": "rattle",
                          "integrator": "lgvc", "dtype": "f64"}}
        log(f"[chains] {args.chain_steps} steps in {el:.3f}s -> {rec['value']:.4e} particle-steps/s")
        snap = d.state()
        if not args.no_cpu:
            cpu = chains_cpu_baseline(snap, p, 15.0)
            if cpu:
                rec["cpu_baseline"] = cpu
                log(f"[chains cpu] reference {cpu['value']:.4e} particle-steps/s ({cpu['seconds']:.1f}s)")
        out["chains_100000"] = rec

        if not args.no_variants:   # the FP32 row
            def chain_forces():
                d.initialize_forces()
                for f in p["forces"]:
                    d.forces(*f)
                return d.get_scalar("POT_ENG")

            err = f32_error(d, snap, chain_forces, FORCE_FIELDS[:3])
            err.pop("per_group")

            def back_to_start():
                d.load_arrays(start)
                for k, v in start_scal.items():
                    d.set_scalar(k, v)
                d.build_nlist()
                d.synchronize()

            # FP32: the same segment (same state, same Langevin counters) after a
            # warm-up that captures the FP32 step graphs
            back_to_start()
            d.run(sx.INT_LGVC, p["forces"], warm, p["dt"], damping=p["gamma"], first_step=step, bonded=True)
            back_to_start()
            rb0 = d.rebuild_count()
            t0 = time.perf_counter()
            d.run(sx.INT_LGVC, p["forces"], args.chain_steps, p["dt"], damping=p["gamma"], first_step=step,
                  bonded=True)
            d.synchronize()
            el = time.perf_counter() - t0
            out["chains_100000_f32"] = f32_row(rec, p["n"], args.chain_steps, el, err, d.rebuild_count() - rb0,
                                               {"T": d.get_scalar("TEMP")})
            log(f"[chains f32] {el / args.chain_steps * 1e3:.4f} ms/step; error {err['force_elem_err']:.2e}")
        d.close()
    if args.batch_steps > 0:
        out["batched_1024"], out["batched_1024_f32"] = batched_workload(sx, args)
    return out


def batched_workload(sx, args):
    """1024 independent LJ / Gay-Berne systems advanced together: one launch
    per step, graph-replayed, launch shape picked by the autotuner."""
    import numpy as np
    from sucrose_amd.workloads import batched_state
    d = sx.System(0)
    p = batched_state(d, log=log)
    step = p["next_step"]
    choice, ms = d.batch_autotune(p["dt"], damping=p["gamma"], first_step=step, nrep=5)
    shape = sx.System.BATCH_SHAPES[choice]
    log(f"[batch] autotune: " + " ".join(f"{s[0]}x{s[1]}:{m:.3f}ms" for s, m in zip(sx.System.BATCH_SHAPES, ms))
        + f" -> {shape[0]} particles per force workgroup, {shape[1]} lanes/particle")
    d.batch_run(50, p["dt"], damping=p["gamma"], first_step=step)
    step += 50
    start, start_step = d.state(), step
    state = start if not args.no_cpu else None
    d.synchronize()
    t0 = time.perf_counter()
    d.batch_run(args.batch_steps, p["dt"], damping=p["gamma"], first_step=step)
    d.synchronize()
    el = time.perf_counter() - t0
    th = d.batch_thermo()
    temps = np.asarray(p["temps"])
    kinds = np.asarray(p["kinds"])
    lj = kinds == sx.LJ
    rec = {"value": p["n"] * args.batch_steps / el, "unit": "particle-steps/s", "steps": args.batch_steps,
           "ms_per_step": el / args.batch_steps * 1e3, "launch_shape": {"particles_per_workgroup": shape[0],
                                                                       "lanes_per_particle": shape[1]},
           "autotune_ms": [round(float(m), 4) for m in ms],
           "T_over_T0": {"lj": float(np.mean(th[lj, 3] / temps[lj])), "gb": float(np.mean(th[~lj, 3] / temps[~lj]))},
           "config": {"workload": "batched_1024", "systems": p["nsys"], "particles": p["n"],
                      "n_per_system": "U[64,512]", "lj": "rho U[0.3,0.8], T U[0.7,1.5], rc min(2.5, L/2)",
                      "gb": "GB(3,5,2,1), rho U[0.08,0.15], T U[2.5,3.5], cutoff min(4, L/2-2)",
                      "integrator": "lgv (gamma 1)", "dt": p["dt"], "pairs": "all pairs I<J per system",
                      "dtype": "f64"}}
    log(f"[batch] {args.batch_steps} steps in {el:.3f}s -> {rec['value']:.4e} particle-steps/s")
    if state is not None:
        cpu = batch_cpu_baseline(state, p)
        if cpu:
            rec["cpu_baseline"] = cpu
            log(f"[batch cpu] reference {cpu['value']:.4e} particle-steps/s ({cpu['seconds']:.1f}s)")
    # FP32 pair terms: the error of one step's forces from one state, then
    # the autotuner, and the timed segment again from its starting state
    step += args.batch_steps
    snap = d.state()

    def batch_step():
        d.batch_run(1, p["dt"], damping=p["gamma"], first_step=step)
        return d.batch_thermo()[:, 2].copy()

    ends = np.cumsum([0] + list(p["counts"]))
    err = f32_error(d, snap, batch_step, groups=[slice(a, b) for a, b in zip(ends[:-1], ends[1:])])
    per = np.asarray(err.pop("per_group"))
    kinds = np.asarray(p["kinds"])
    err["force_elem_err_lj"] = float(per[kinds == sx.LJ].max())
    err["force_elem_err_gb"] = float(per[kinds == sx.GB].max())
    err["systems_over_1e-5"] = int(np.sum(per > 1e-5))
    choice, ms = d.batch_autotune(p["dt"], damping=p["gamma"], first_step=step, nrep=5)
    shape = sx.System.BATCH_SHAPES[choice]
    d.load_arrays(start)
    d.batch_run(50, p["dt"], damping=p["gamma"], first_step=start_step)   # captures the step graph
    d.load_arrays(start)
    d.synchronize()
    t0 = time.perf_counter()
    d.batch_run(args.batch_steps, p["dt"], damping=p["gamma"], first_step=start_step)
    d.synchronize()
    el = time.perf_counter() - t0
    row = f32_row(rec, p["n"], args.batch_steps, el, err, 0,
                  {"launch_shape": {"particles_per_workgroup": shape[0], "lanes_per_particle": shape[1]},
                   "autotune_ms": [round(float(m), 4) for m in ms]})
    del row["rebuilds"]
    log(f"[batch f32] {el / args.batch_steps * 1e3:.4f} ms/step (shape {shape}); error {err['force_elem_err']:.2e} "
        f"(LJ {err['force_elem_err_lj']:.2e}, GB {err['force_elem_err_gb']:.2e}, "
        f"{err['systems_over_1e-5']} systems over 1e-5)")
    d.close()
    return rec, row


def batch_cpu_baseline(state, p, nsample=8, seconds_each=1.5):
    """The reference's own Langevin loop (Verlet list, skin 0.3, its rebuild
    rule) run system by system on evenly spaced systems of the batch, each in
    a private copy of the reference library."""
    from oracle.oracle import FIELDS, GB, LGV, LJ, Reference, State, reference_available
    if not reference_available(fast=True):
        return None
    counts = p["counts"]
    off = [0]
    for c in counts:
        off.append(off[-1] + c)
    picks = [int(k) for k in range(0, p["nsys"], max(1, p["nsys"] // nsample))][:nsample]
    tot_ps, tot_t = 0.0, 0.0
    for s in picks:
        n, L, kind, prm = counts[s], p["boxes"][s], p["kinds"][s], p["params"][s]
        st = State(n, [L, L, L])
        for k, f in enumerate(FIELDS):
            st.buf[:, k] = state[f.upper()][off[s]:off[s] + n]
        st["target_temp"] = p["temps"][s]
        st["dof"] = 6 * n - 3
        st["density"] = n / L ** 3
        R = Reference(fast=True, private=True)
        R.load(st)
        if kind == LJ:
            fk, fp, reach = LJ, prm[:3], prm[2]
        else:
            fk, fp, reach = GB, [prm[0]], prm[0] + 2.0
        R.nlist_init(0.3, reach)
        t0 = time.time()
        R.run_steps(LGV, fk, fp, 1, 1, p["dt"], p["gamma"])
        t1 = max(time.time() - t0, 1e-5)
        k = int(max(1, min(2000, seconds_each / t1)))
        t0 = time.time()
        R.run_steps(LGV, fk, fp, k, 2, p["dt"], p["gamma"])
        tot_t += time.time() - t0
        tot_ps += n * k
    return {"value": tot_ps / tot_t, "unit": "particle-steps/s", "cores": 1, "kind": "reference",
            "sample": f"the reference loop (lgv, lj or gb pair loop over its Verlet list) on {len(picks)} evenly "
                      f"spaced systems of the batch, one after another, ~{seconds_each}s each",
            "seconds": tot_t}


TRAIN_SCHEDULE = [(3e-3, 10000), (1.5e-3, 10000), (7e-4, 10000), (3e-4, 6000), (1e-4, 4000)]


def learned_vs_plain(dev, sx, p, step, nsteps):
    """The lambda* network trained on the GPU by the library's own kernels
    on pairs the contact solver generates along this trajectory, then the
    converged contact solve with and without its warm start: each solver
    timed alone on the same work list, forces compared, and the whole loop
    timed in each mode on the same trajectory start."""
    import numpy as np
    n, dt, gamma = p["n"], p["dt"], p["gamma"]
    t0 = time.perf_counter()
    dev.mlp_init(seed=2024)
    for _ in range(48):       # ~2M pairs: 48 work lists, 10 steps apart
        dev.run(sx.INT_LGV, [(sx.ECPR,)], 10, dt, damping=gamma, first_step=step)
        step += 10
        samples = dev.mlp_collect(sx.ECPR, capacity=1 << 21)
    t_collect = time.perf_counter() - t0
    t0 = time.perf_counter()
    losses = []
    for k, (lr, nst) in enumerate(TRAIN_SCHEDULE):
        tr = dev.mlp_train(nst, batch=4096, lr=lr, w=(1.0, 0.1), seed=100 + k)
        losses.append((float(tr[:50].mean()), float(tr[-200:].mean())))
    dev.synchronize()
    t_train = time.perf_counter() - t0
    train_steps = sum(x[1] for x in TRAIN_SCHEDULE)
    log(f"[learned] {samples} samples in {t_collect:.2f}s; {train_steps} Adam steps (batch 4096) in {t_train:.2f}s; "
        f"loss {losses[0][0]:.3e} -> {losses[-1][1]:.3e}")
    dev.run(sx.INT_LGV, [(sx.ECPR,)], 10, dt, damping=gamma, first_step=step)
    step += 10
    cmp = dev.compare_solvers(sx.ECPR, nrep=20)
    # forces and torques on one state: reference Brent, plain converged, learned
    ft = {}
    for name, solver in (("brent", sx.ECP_BRENT_REFERENCE), ("plain", sx.ECP_CONVERGED),
                         ("learned", sx.ECP_LEARNED)):
        dev.set_ecp_solver(solver)
        dev.initialize_forces()
        dev.forces(sx.ECPR)
        ft[name] = np.stack([dev.download(f) for f in ("FX", "FY", "FZ", "TX", "TY", "TZ")])
    scale = np.abs(ft["plain"]).max()
    err_learned = float(np.abs(ft["learned"] - ft["plain"]).max() / scale)
    err_brent = float(np.abs(ft["brent"] - ft["plain"]).max() / scale)
    # the whole loop in each mode from the same state
    state = dev.state()
    rates = {}
    for name, solver in (("plain", sx.ECP_CONVERGED), ("learned", sx.ECP_LEARNED)):
        for f, v in state.items():
            dev.upload(f, v)
        dev.build_nlist()
        dev.set_ecp_solver(solver)
        dev.run(sx.INT_LGV, [(sx.ECPR,)], 50, dt, damping=gamma, first_step=step)
        dev.synchronize()
        t0 = time.perf_counter()
        dev.run(sx.INT_LGV, [(sx.ECPR,)], nsteps, dt, damping=gamma, first_step=step + 50)
        dev.synchronize()
        rates[name] = n * nsteps / (time.perf_counter() - t0)
    dev.set_ecp_solver(sx.ECP_BRENT_REFERENCE)
    for f, v in state.items():
        dev.upload(f, v)
    dev.build_nlist()
    lam_plain = cmp["ms_newton_plain"]
    lam_learned = cmp["ms_network"] + cmp["ms_newton_warm"]
    # Amdahl bound of any warm start on this stage, from the same timings: the
    # time of one Newton round of a wavefront (the difference between the
    # plain and warm-started lambda solves over the difference of their
    # per-wavefront maximum iterations), the part of the plain stage those
    # rounds take, and the stage time left with a free network and (a) no
    # iterations at all, (b) the warm start's own rounds
    d_it = cmp["wave_iters_plain"] - cmp["wave_iters_warm"]
    t_round = (lam_plain - cmp["ms_newton_warm"]) / d_it if d_it > 0 else float("nan")
    removable = t_round * cmp["wave_iters_plain"]
    floor = cmp["ms_solve_plain"] - removable
    amdahl = {"ms_per_newton_round": t_round, "ms_iterations_in_plain_stage": removable,
              "ms_stage_without_iterations": floor,
              "max_speedup_free_network_no_iterations": cmp["ms_solve_plain"] / floor if floor > 0 else None,
              "max_speedup_free_network_warm_rounds": cmp["ms_solve_plain"] / (floor + t_round * cmp["wave_iters_warm"])}
    out = {
        "training": {"samples": samples, "adam_steps": train_steps, "batch": 4096, "seconds": t_train,
                     "loss_first": losses[0][0], "loss_last": losses[-1][1], "dtype": "bf16 mfma, fp32 accumulate"},
        "lambda0_err_rms": cmp["lambda0_err_rms"], "lambda0_err_max": cmp["lambda0_err_max"],
        "pairs": cmp["pairs"], "proven_out_of_range_by_warm_start": cmp["proven_out_of_range"],
        "newton_iters_per_pair": {"plain": cmp["iters_plain"], "warm": cmp["iters_warm"]},
        "wave_max_iters_per_round": {"plain": cmp["wave_iters_plain"], "warm": cmp["wave_iters_warm"],
                                     "brent_1e-10": cmp["wave_iters_brent_tight"]},
        "lambda_solve_ms": {"brent_reference_tol": cmp["ms_brent"], "brent_1e-10": cmp["ms_brent_tight"],
                            "newton_plain": lam_plain, "network": cmp["ms_network"],
                            "newton_warm": cmp["ms_newton_warm"], "learned_total": lam_learned},
        "speedup_vs_plain_newton": lam_plain / lam_learned,
        "speedup_vs_brent_1e-10": cmp["ms_brent_tight"] / lam_learned,
        "solve_stage_ms": {"plain": cmp["ms_solve_plain"], "learned": cmp["ms_solve_learned"]},
        "solve_stage_speedup_vs_plain": cmp["ms_solve_plain"] / cmp["ms_solve_learned"],
        "amdahl_bound": amdahl,
        "force_rel_err_learned_vs_plain": err_learned,
        "force_rel_err_brent_reference_vs_converged": err_brent,
        "step_particle_steps_per_s": rates,
        "_next_step": step + 50 + nsteps,
    }
    log(f"[learned] lambda0 rms {cmp['lambda0_err_rms']:.2e}; lambda solve ms plain {lam_plain:.4f} learned "
        f"{lam_learned:.4f} (network {cmp['ms_network']:.4f}) brent-1e-10 {cmp['ms_brent_tight']:.4f}; "
        f"solve stage ms plain {cmp['ms_solve_plain']:.4f} learned (one kernel) {cmp['ms_solve_learned']:.4f}; "
        f"force err learned {err_learned:.1e}, reference-brent {err_brent:.1e}; steps/s {rates}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-plain contact solver section")
    ap.add_argument("--no-others", action="store_true", help="skip the Gay-Berne and chain workloads")
    ap.add_argument("--gb-steps", type=int, default=2000)
    ap.add_argument("--chain-steps", type=int, default=1000)
    ap.add_argument("--batch-steps", type=int, default=5000, help="steps of the batched replicas (0 skips them)")
    ap.add_argument("--learned-steps", type=int, default=1000,
                    help="steps of the whole loop timed for each contact solver")
    ap.add_argument("--save-state", default=None, help="write the prepared workload state to this .npz")
    ap.add_argument("--save-states", default=None,
                    help="write the prepared Gay-Berne and charged-chain workloads into this directory")
    ap.add_argument("--load-states", default=None,
                    help="start the Gay-Berne and charged-chain workloads from a --save-states directory")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the FP32 rows and the list-radius variant of the other workloads")
    ap.add_argument("--load-state", default=None,
                    help="start from a state written by --save-state instead of preparing the workload")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device copy 10 times (reference point for HBM counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4, "copies": 10}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported (this build runs one "
              f"process on one MI355X; multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    import sucrose_amd as sx
    from sucrose_amd.workloads import ecp_langevin_state, load_state, save_state

    dev = sx.System(0)
    if args.load_state:
        p = load_state(dev, args.load_state)
        log(f"[workload] loaded {args.load_state}")
    else:
        p = ecp_langevin_state(dev, n=args.n, log=log)
    if args.save_state:
        save_state(dev, p, args.save_state)
    n, dt, gamma = p["n"], p["dt"], p["gamma"]
    forces = [(sx.ECPR,)]
    step = p["next_step"]
    dev.run(sx.INT_LGV, forces, args.warmup, dt, damping=gamma, first_step=step)
    step += args.warmup
    dev.synchronize()
    rb0 = dev.rebuild_count()
    t0 = time.perf_counter()
    dev.run(sx.INT_LGV, forces, args.steps, dt, damping=gamma, first_step=step)
    dev.synchronize()
    el = time.perf_counter() - t0
    step += args.steps
    rebuilds = dev.rebuild_count() - rb0
    value = n * args.steps / el
    temp = dev.get_scalar("TEMP")
    pe = dev.get_scalar("POT_ENG")
    log(f"[gpu] {args.steps} steps in {el:.3f}s -> {value:.4e} particle-steps/s; T={temp:.4f} PE/N={pe / n:.5f} "
        f"rebuilds={rebuilds}")

    # roofline of the dominant kernel, timed live with HIP events on the library stream
    ms, st = dev.profile_ecp(sx.ECPR, (), nrep=20)
    screened, solved, iters = st
    flops = {
        "ecp_precompute": n * 150.0,
        "ecp_screen": screened * FLOP_SCREEN,
        "ecp_solve": solved * (30 + FLOP_OPTIM_DIST + FLOP_ECP_FORCE) + iters * (FLOP_OPTIM_DIST + FLOP_BRENT_ITER),
        "ecp_apply": 2 * solved * FLOP_APPLY_ENTRY,
    }
    names = list(flops)
    k = max(range(4), key=lambda q: ms[q])
    ach = flops[names[k]] / (ms[k] * 1e-3) / 1e12
    issued = ISSUED_FP64.get(names[k])
    roof = {"bound": "fp64-valu", "achieved": ach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": ach / PEAK_FP64_TFLOPS, "traffic": TRAFFIC_BYTES.get(names[k]), "kernel": names[k],
            "traffic_unit": "bytes per launch (L2 misses: rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated)",
            "peak_kind": "fp64 vector ALU (MI355X fp64 dense figure); no matrix-core work in this kernel",
            "issued_fp64_flops_per_launch": issued,
            "issued_achieved": issued / (ms[k] * 1e-3) / 1e12 if issued else None,
            "issued_note": "rocprofv3 SQ_INSTS_VALU_FLOPS_FP64 x 64 lanes: every issued FP64 operation of the "
                           "kernel, inactive lanes and the division/square-root refinement steps included",
            "stage_ms": {names[q]: ms[q] for q in range(4)},
            "pairs_screened": screened, "pairs_solved": solved, "brent_iters_per_solved_pair": iters / max(solved, 1)}
    log(f"[gpu] stages ms {roof['stage_ms']} screened={screened:.0f} solved={solved:.0f} "
        f"iters/pair={iters / max(solved, 1):.2f}")

    learned = None if args.no_learned else learned_vs_plain(dev, sx, p, step, args.learned_steps)
    if learned:
        step = learned.pop("_next_step")

    others = None if args.no_others else other_workloads(sx, args)

    cpu = None
    if not args.no_cpu:
        dev.initialize_forces()
        dev.forces(sx.ECPR)
        arrays = dev.state()
        p["next_step"] = step
        cpu = cpu_baseline(arrays, p["box"], p, args.cpu_seconds)
        cpu["cores"] = 1
        cpu["cpu_model"], cpu["host_cpus"] = host_cpu()
        log(f"[cpu] {cpu['kind']} {cpu['value']:.4e} particle-steps/s ({cpu['seconds']:.1f}s) on {cpu['cpu_model']}")
        allc = cpu_all_cores(arrays, p["box"], p, min(10.0, args.cpu_seconds))
        if allc:
            cpu["all_cores"] = allc
            log(f"[cpu] all cores ({allc['cores']} threads) {allc['value']:.4e} particle-steps/s")

    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "ecp_langevin_32768", "n": n, "phi": p["phi"], "semi_axes": [0.5, 0.5, 1.5],
                   "T": p["T"], "gamma": gamma, "dt": dt, "cutoff": p["cutoff"], "skin": p["skin"],
                   "potential": "ecpr", "lambda_solver": "brent_reference", "rebuild_rule": "half_skin",
                   "parallelism": "single_gpu"},
        "roofline": roof,
        "learned_vs_plain": learned,
        "other_workloads": others,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(out), flush=True)
    dev.close()


if __name__ == "__main__":
    main()
