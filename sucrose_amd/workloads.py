"""Seeded synthetic workloads of the benchmark configurations, prepared on the
device with the library's own routines.

GB 8000 (Gay-Berne GB(3,5,2,1) with a Nose-Hoover chain): 20^3 particles on a
cubic lattice with uniformly random orientations.  At number density 0.30
(lattice spacing 1.494) randomly oriented 3:1 ellipsoids overlap, so the
lattice starts at spacing 3.2 (no overlap for any orientation) and is
compressed affinely to spacing 1.494 in stages interleaved with short
Langevin runs of the Gay-Berne potential itself, then equilibrated with the
Nose-Hoover chain.

Chains 100000: 2000 self-avoiding random-walk chains of 50 beads (bond 1.0,
no two beads closer than 0.95) at bead density 0.5, charges alternating +1/-1
along each chain, RATTLE bonds, cosine bend k=20 about the straight chain,
dihedral k(1 + cos phi) with k=1, WCA (sigma=1) and Debye-Hueckel (kappa=1,
cutoff 5, dielectric 1) over one neighbour list.

Batched replicas: nsys independent systems of N ~ U[64, 512] particles, even
indices Lennard-Jones and odd ones Gay-Berne GB(3,5,2,1), each in its own
cubic periodic box.  LJ systems: number density U[0.3, 0.8], T U[0.7, 1.5],
cutoff min(2.5, L/2), a simple cubic lattice of ceil(N^(1/3))^3 sites with N
of them occupied at random and a small jitter.  GB systems: T U[2.5, 3.5] and
number density U[0.08, 0.15] (a 3:1 ellipsoid reaches its neighbours' images
through the minimum image only if the interaction range, cutoff + 2, stays
within L/2; at N = 64 and density 0.3 that leaves a cutoff below the
potential's minimum), cutoff min(4, L/2 - 2), random sequential addition of
hard (0.5, 0.5, 1.5) ellipsoids.  Velocities and angular momenta are drawn
per system from the Maxwell-Boltzmann distribution at that system's
temperature, then all systems are equilibrated together by the batched
Langevin step itself.

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

from . import (DH, ECPR, GB, INT_LGV, LJ, INT_LGVC, INT_NHT, REBUILD_HALF_SKIN, SCALARS, WCA, System,
               count_ellipsoid_overlaps, gen_chains, gen_ellipsoids_rsa, uniform_quaternions)

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


GB_SIG = (1.0, 1.0, 3.0)   # kappa = 3
GB_EPS = (5.0, 5.0, 1.0)   # kappa' = 5


def gb_nhc_state(dev: System, n_side=20, rho=0.30, seed=15, T=3.0, dt=0.0015, cutoff=4.0, skin=0.4, mu=2, nu=1,
                 a0=3.2, stages=60, steps_per_stage=30, dt_compress=0.001, equil_steps=200, log=print):
    """Builds the Gay-Berne + Nose-Hoover workload on dev; returns its parameters."""
    t0 = time.time()
    n = n_side ** 3
    L1 = (n / rho) ** (1.0 / 3.0)
    L0 = a0 * n_side
    g = (np.arange(n_side) + 0.5) * a0 - L0 / 2
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    quat = uniform_quaternions(n, seed)
    dev.set_num_atoms(n)
    dev.set_box(L0)
    for k, f in enumerate(("RX", "RY", "RZ")):
        dev.upload(f, pos[:, k])
    for k, f in enumerate(("QW", "QX", "QY", "QZ")):
        dev.upload(f, quat[:, k])
    for k, ax in enumerate("XYZ"):
        dev.upload("SIG_" + ax, GB_SIG[k])
        dev.upload("EPS_" + ax, GB_EPS[k])
    dev.set_target_temperature(T)
    dev.initialize_system()
    dev.set_velocity_to_temperature(T, seed)
    dev.set_rebuild_rule(REBUILD_HALF_SKIN)
    dev.init_nlist(skin, cutoff)
    gb = [(GB, cutoff, mu, nu)]
    step = 1
    Lc = L0
    for s in range(1, stages + 1):
        Ln = L0 * (L1 / L0) ** (s / stages)
        f = Ln / Lc
        for c in ("RX", "RY", "RZ"):
            dev.upload(c, dev.download(c) * f)
        dev.set_box(Ln)
        dev.init_nlist(skin, cutoff)
        dev.run(INT_LGV, gb, steps_per_stage, dt_compress, damping=1.0, first_step=step)
        step += steps_per_stage
        Lc = Ln
    dev.run(INT_NHT, gb, equil_steps, dt, first_step=step)
    step += equil_steps
    pe = dev.get_scalar("POT_ENG")
    if not np.isfinite(pe):
        raise RuntimeError("Gay-Berne workload preparation diverged")
    log(f"[workload] GB {n}: lattice {a0} -> {L1 / n_side:.4f} spacing (L={L1:.4f}), NHC equilibrated, "
        f"PE/N={pe / n:.4f} T={dev.get_scalar('TEMP'):.4f}, {time.time() - t0:.1f}s")
    return dict(n=n, box=L1, rho=rho, T=T, dt=dt, cutoff=cutoff, skin=skin, mu=mu, nu=nu, next_step=step, seed=seed)


def chains_state(dev: System, nchains=2000, nbeads=50, rho=0.5, seed=7, T=1.0, dt=0.005, gamma=1.0, bond=1.0,
                 min_dist=0.95, k_bend=20.0, bend_angle=180.0, k_dih=1.0, dh=(5.0, 1.0, 1.0), skin=0.8,
                 equil_steps=200, log=print):
    """Builds the charged-chain workload on dev; returns its parameters and the force list."""
    t0 = time.time()
    n = nchains * nbeads
    L = (n / rho) ** (1.0 / 3.0)
    pos = gen_chains(nchains, nbeads, L, bond=bond, min_dist=min_dist, seed=seed)
    dev.set_num_atoms(n)
    dev.set_box(L)
    for k, f in enumerate(("RX", "RY", "RZ")):
        dev.upload(f, pos[:, k])
    charge = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    dev.upload("CHARGE", charge)
    chains_topology(dev, nchains, nbeads, bond, k_bend, bend_angle, k_dih, bonded=False)
    dev.set_target_temperature(T)
    dev.initialize_system()
    dev.set_velocity_to_temperature(T, seed)
    chains_topology(dev, nchains, nbeads, bond, k_bend, bend_angle, k_dih, constraints=False)
    dev.set_rebuild_rule(REBUILD_HALF_SKIN)
    dev.init_nlist(skin, dh[0])
    forces = [(WCA, 1.0, 1.0), (DH,) + tuple(dh)]
    dev.run(INT_LGVC, forces, equil_steps, dt, damping=gamma, first_step=1, bonded=True)
    pe = dev.get_scalar("POT_ENG")
    if not np.isfinite(pe):
        raise RuntimeError("chain workload preparation diverged")
    log(f"[workload] chains {nchains}x{nbeads} (L={L:.4f}) equilibrated {equil_steps} steps, PE/N={pe / n:.4f} "
        f"T={dev.get_scalar('TEMP'):.4f}, {time.time() - t0:.1f}s")
    return dict(n=n, box=L, rho=rho, T=T, dt=dt, gamma=gamma, skin=skin, cutoff=dh[0], next_step=1 + equil_steps,
                seed=seed, forces=forces, nchains=nchains, nbeads=nbeads, bond=bond, k_bend=k_bend,
                bend_angle=bend_angle, k_dih=k_dih)


def chains_topology(dev: System, nchains, nbeads, bond, k_bend, bend_angle, k_dih, constraints=True, bonded=True):
    """The chains' RATTLE bonds (consecutive beads) and bonded terms: cosine
    bend about the straight chain (cosine_bend_, force.f90:2345) and the
    dihedral k(1 + cos phi) (dih_, :2286)."""
    ids = np.arange(nchains * nbeads).reshape(nchains, nbeads) + 1          # 1-based
    if constraints:
        bi = ids[:, :-1].ravel()
        bj = ids[:, 1:].ravel()
        dev.set_constraints(bi, bj, np.full(len(bi), bond))
    if bonded:
        dev.set_bonded(2, np.stack([ids[:, :-2], ids[:, 1:-1], ids[:, 2:]], -1).reshape(-1), k_bend, bend_angle)
        dev.set_bonded(3, np.stack([ids[:, :-3], ids[:, 1:-2], ids[:, 2:-1], ids[:, 3:]], -1).reshape(-1), k_dih)


def batched_layout(nsys=1024, nmin=64, nmax=512, seed=31):
    """Host arrays of the batched workload: per-system counts, boxes, kinds,
    parameters (LJ: sigma, epsilon, cutoff; GB: cutoff, mu, nu) and
    temperatures, and per-particle positions (centred boxes), quaternions and
    sig/eps axes in system order."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(nmin, nmax + 1, nsys).astype(np.int32)
    kinds = np.where(np.arange(nsys) % 2 == 0, LJ, GB).astype(np.int32)
    boxes = np.zeros(nsys)
    params = np.zeros((nsys, 4))
    temps = np.zeros(nsys)
    pos, quat, sig, eps = [], [], [], []
    for s in range(nsys):
        n = int(counts[s])
        if kinds[s] == LJ:
            rho = rng.uniform(0.3, 0.8)
            L = (n / rho) ** (1.0 / 3.0)
            m = int(math.ceil(n ** (1.0 / 3.0) - 1e-9))
            a = L / m
            g = (np.arange(m) + 0.5) * a - L / 2
            sites = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
            p = sites[np.sort(rng.permutation(len(sites))[:n])] + rng.uniform(-0.05, 0.05, (n, 3)) * a
            q = np.zeros((n, 4)); q[:, 0] = 1.0
            params[s, :3] = (1.0, 1.0, min(2.5, 0.5 * L - 1e-9))
            temps[s] = rng.uniform(0.7, 1.5)
            sig.append(np.ones((n, 3))); eps.append(np.ones((n, 3)))
        else:
            rho = rng.uniform(0.08, 0.15)
            L = (n / rho) ** (1.0 / 3.0)
            p, q = gen_ellipsoids_rsa(n, ELLIPSOID_SEMI_AXES, L, int(rng.integers(1, 2 ** 31)))
            if len(p) != n:
                raise RuntimeError(f"placed {len(p)} of {n} ellipsoids")
            # interaction range cutoff + sig_z - sigma_0 = cutoff + 2 within L/2 (with a rounding margin)
            params[s, :3] = (min(4.0, 0.5 * L - 2.0 - 1e-9), 2, 1)
            temps[s] = rng.uniform(2.5, 3.5)
            sig.append(np.tile(GB_SIG, (n, 1)).astype(np.float64)); eps.append(np.tile(GB_EPS, (n, 1)).astype(float))
        boxes[s] = L
        pos.append(p); quat.append(q)
    return dict(counts=counts, boxes=boxes, kinds=kinds, params=params, temps=temps, pos=np.concatenate(pos),
                quat=np.concatenate(quat), sig=np.concatenate(sig), eps=np.concatenate(eps))


def _rot(q):
    """R(q) of each unit quaternion (w x y z): space = R body."""
    w, x, y, z = q.T
    return np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
                     np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
                     np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], -2)


def batched_state(dev: System, nsys=1024, nmin=64, nmax=512, seed=31, dt=0.0015, gamma=1.0, relax_steps=100,
                  equil_steps=400, log=print):
    """Builds the batched-replica workload on dev and equilibrates it with the
    batched Langevin step; returns its parameters."""
    t0 = time.time()
    lay = batched_layout(nsys, nmin, nmax, seed)
    counts = lay["counts"]
    n = int(counts.sum())
    dev.set_num_atoms(n)
    dev.set_box(float(lay["boxes"].max()))
    for k, f in enumerate(("RX", "RY", "RZ")):
        dev.upload(f, lay["pos"][:, k])
    for k, f in enumerate(("QW", "QX", "QY", "QZ")):
        dev.upload(f, lay["quat"][:, k])
    for k, ax in enumerate("XYZ"):
        dev.upload("SIG_" + ax, lay["sig"][:, k])
        dev.upload("EPS_" + ax, lay["eps"][:, k])
    dev.set_target_temperature(1.0)
    dev.initialize_system()
    rng = np.random.default_rng(seed + 1)
    T = np.repeat(lay["temps"], counts)
    mass = dev.download("MASS")
    v = rng.normal(size=(n, 3)) * np.sqrt(T / mass)[:, None]
    sys_id = np.repeat(np.arange(nsys), counts)
    for k in range(3):   # no centre-of-mass drift within a system
        mom = np.bincount(sys_id, weights=mass * v[:, k], minlength=nsys)
        v[:, k] -= (mom / np.bincount(sys_id, weights=mass, minlength=nsys))[sys_id]
    inertia = np.stack([dev.download(f) for f in ("IXX", "IYY", "IZZ")], 1)
    lb = rng.normal(size=(n, 3)) * np.sqrt(T[:, None] * inertia)
    ls = np.einsum("nij,nj->ni", _rot(lay["quat"]), lb)
    for k, f in enumerate(("VX", "VY", "VZ")):
        dev.upload(f, v[:, k])
    for k, f in enumerate(("LX", "LY", "LZ")):
        dev.upload(f, ls[:, k])
    dev.batch_set(counts, lay["boxes"], lay["kinds"], lay["params"], lay["temps"])
    step = 1
    dev.batch_run(relax_steps, dt / 5, damping=gamma, first_step=step)
    step += relax_steps
    dev.batch_run(equil_steps, dt, damping=gamma, first_step=step)
    step += equil_steps
    th = dev.batch_thermo()
    if not np.all(np.isfinite(th)):
        raise RuntimeError("batched workload preparation diverged")
    lj = lay["kinds"] == LJ
    log(f"[workload] batch {nsys} systems, {n} particles: equilibrated {relax_steps}+{equil_steps} steps, "
        f"T/T0 LJ {np.mean(th[lj, 3] / lay['temps'][lj]):.3f} GB {np.mean(th[~lj, 3] / lay['temps'][~lj]):.3f}, "
        f"{time.time() - t0:.1f}s")
    return dict(n=n, nsys=nsys, dt=dt, gamma=gamma, next_step=step, seed=seed, counts=counts.tolist(),
                boxes=lay["boxes"].tolist(), kinds=lay["kinds"].tolist(), params=lay["params"].tolist(),
                temps=lay["temps"].tolist())


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


def save_prepared(dev: System, params: dict, path):
    """Writes a prepared workload (device arrays, scalars, parameters) to an
    .npz file, so that a later run starts from it instead of preparing it."""
    arrays = {f"f_{k}": v for k, v in dev.state().items()}
    scal = {k: dev.get_scalar(k) for k in SCALARS}
    np.savez(path, params=np.array(json.dumps(params)), scal=np.array(json.dumps(scal)), **arrays)


def load_prepared(dev: System, path):
    """Restores a workload written by save_prepared (for the charged chains
    also their bonds and bonded terms) and initialises its neighbour list;
    returns the workload parameters."""
    z = np.load(path, allow_pickle=False)
    params = json.loads(str(z["params"]))
    scal = json.loads(str(z["scal"]))
    dev.set_num_atoms(params["n"])
    dev.set_box(params["box"])
    for key in z.files:
        if key.startswith("f_"):
            dev.upload(key[2:], z[key])
    if "nchains" in params:
        chains_topology(dev, params["nchains"], params["nbeads"], params["bond"], params["k_bend"],
                        params["bend_angle"], params["k_dih"])
    dev.set_target_temperature(params["T"])
    dev.initialize_system(remove_com=False)
    for key in z.files:   # initialize_system derives sums only; the fields as saved
        if key.startswith("f_"):
            dev.upload(key[2:], z[key])
    for k, v in scal.items():
        dev.set_scalar(k, v)
    dev.set_rebuild_rule(REBUILD_HALF_SKIN)
    dev.init_nlist(params["skin"], params["cutoff"])
    if "forces" in params:
        params["forces"] = [tuple(f) for f in params["forces"]]
    return params


def overlaps_on_device(dev: System, box):
    pos = np.stack([dev.download(c) for c in ("RX", "RY", "RZ")], 1)
    quat = np.stack([dev.download(c) for c in ("QW", "QX", "QY", "QZ")], 1)
    return count_ellipsoid_overlaps(pos, quat, ELLIPSOID_SEMI_AXES, box)
