"""Seeded synthetic workloads of the benchmark configurations, prepared on the
device with the library's own routines.

ECP 32k (the headline): identical prolate ellipsoids, semi-axes
(0.5, 0.5, 1.5) (sig = 1, 1, 3), N = 32768, packing fraction 0.35, random
positions and orientations with no overlaps.  Random sequential addition
jams near phi = 0.31 for this shape, so the state is made by RSA at
phi0 = 0.25 followed by a gradual affine compression to 0.35 interleaved
with short Langevin runs of the repulsive ECP potential itself (each stage
shrinks lengths by ~0.2 %, far less than the potential's range, so no stage
creates a deep overlap), then an equilibration run at the final density.
"""
from __future__ import annotations

import json
import math
import time

import numpy as np

from . import ECPR, INT_LGV, REBUILD_HALF_SKIN, System, count_ellipsoid_overlaps, gen_ellipsoids_rsa

ELLIPSOID_SEMI_AXES = (0.5, 0.5, 1.5)


def ellipsoid_volume(semi=ELLIPSOID_SEMI_AXES):
    return 4.0 / 3.0 * math.pi * semi[0] * semi[1] * semi[2]


def ecp_langevin_state(dev: System, n=32768, phi=0.35, phi0=0.25, seed=2024, T=1.0, gamma=1.0, dt=0.002,
                       cutoff=3.1, skin=0.3, stages=40, steps_per_stage=50, equil_steps=500, log=print):
    """Builds the ECP Langevin workload on dev; returns its parameters."""
    t0 = time.time()
    v = ellipsoid_volume()
    L0 = (n * v / phi0) ** (1.0 / 3.0)
    L1 = (n * v / phi) ** (1.0 / 3.0)
    pos, quat = gen_ellipsoids_rsa(n, ELLIPSOID_SEMI_AXES, L0, seed)
    if len(pos) != n:
        raise RuntimeError(f"RSA placed {len(pos)} of {n} ellipsoids")
    log(f"[workload] RSA phi0={phi0} placed {n} in {time.time() - t0:.1f}s")
    dev.set_num_atoms(n)
    dev.set_box(L0)
    for k, f in enumerate(("RX", "RY", "RZ")):
        dev.upload(f, pos[:, k])
    for k, f in enumerate(("QW", "QX", "QY", "QZ")):
        dev.upload(f, quat[:, k])
    dev.upload("SIG_X", 2 * ELLIPSOID_SEMI_AXES[0])
    dev.upload("SIG_Y", 2 * ELLIPSOID_SEMI_AXES[1])
    dev.upload("SIG_Z", 2 * ELLIPSOID_SEMI_AXES[2])
    dev.set_target_temperature(T)
    dev.initialize_system()
    dev.set_velocity_to_temperature(T, seed)
    dev.set_rebuild_rule(REBUILD_HALF_SKIN)
    dev.init_nlist(skin, cutoff)
    step = 1
    Lc = L0
    for s in range(1, stages + 1):
        Ln = L0 * (L1 / L0) ** (s / stages)
        f = Ln / Lc
        for c in ("RX", "RY", "RZ"):
            dev.upload(c, dev.download(c) * f)
        dev.set_box(Ln)
        dev.init_nlist(skin, cutoff)
        dev.run(INT_LGV, [(ECPR,)], steps_per_stage, dt, damping=gamma, first_step=step)
        step += steps_per_stage
        Lc = Ln
    dev.run(INT_LGV, [(ECPR,)], equil_steps, dt, damping=gamma, first_step=step)
    step += equil_steps
    log(f"[workload] compressed to phi={phi} (L={L1:.4f}) and equilibrated, {time.time() - t0:.1f}s total")
    return dict(n=n, box=L1, phi=phi, T=T, gamma=gamma, dt=dt, cutoff=cutoff, skin=skin, next_step=step,
                seed=seed)


def save_state(dev: System, params: dict, path):
    """Writes the device state and the workload parameters to an .npz file."""
    arrays = {f"f_{k}": v for k, v in dev.state().items()}
    np.savez(path, params=np.array(json.dumps(params)), **arrays)


def load_state(dev: System, path):
    """Restores a state written by save_state (same particle count and box),
    re-initialises the neighbour list; returns the workload parameters."""
    z = np.load(path, allow_pickle=False)
    params = json.loads(str(z["params"]))
    dev.set_num_atoms(params["n"])
    dev.set_box(params["box"])
    for key in z.files:
        if key.startswith("f_"):
            dev.upload(key[2:], z[key])
    dev.set_target_temperature(params["T"])
    dev.initialize_system(remove_com=False)
    for key in z.files:   # initialize_system resets velocities' derived sums only; restore fields again
        if key.startswith("f_"):
            dev.upload(key[2:], z[key])
    dev.set_rebuild_rule(REBUILD_HALF_SKIN)
    dev.init_nlist(params["skin"], params["cutoff"])
    return params


def overlaps_on_device(dev: System, box):
    pos = np.stack([dev.download(c) for c in ("RX", "RY", "RZ")], 1)
    quat = np.stack([dev.download(c) for c in ("QW", "QX", "QY", "QZ")], 1)
    return count_ellipsoid_overlaps(pos, quat, ELLIPSOID_SEMI_AXES, box)
