"""Seeded synthetic states shared by the tests (host FP64, oracle layout)."""
from __future__ import annotations

import numpy as np

from oracle.oracle import FIELDS, State


def rand_quats(rng, n):
    q = rng.normal(size=(n, 4))
    return q / np.linalg.norm(q, axis=1)[:, None]


def base_state(rng, n, box, sig=(1.0, 1.0, 1.0), eps=(1.0, 1.0, 1.0), mass=(0.5, 2.0), T=1.5,
               forces=True, charges=False):
    box = np.broadcast_to(np.asarray(box, dtype=np.float64), (3,)).copy()
    st = State(n, box)
    for k, f in enumerate(["rx", "ry", "rz"]):
        st[f] = rng.uniform(-box[k] / 2, box[k] / 2, n)
    for f in ["vx", "vy", "vz", "lx", "ly", "lz"]:
        st[f] = rng.normal(size=n)
    if forces:
        for f in ["fx", "fy", "fz", "tx", "ty", "tz"]:
            st[f] = rng.normal(size=n) * 5.0
    q = rand_quats(rng, n)
    for k, f in enumerate(["qw", "qx", "qy", "qz"]):
        st[f] = q[:, k]
    st["mass"] = rng.uniform(mass[0], mass[1], n)
    st["inv_mass"] = 1.0 / st["mass"]
    st["charge"] = rng.choice([-1.0, 1.0], n) if charges else 0.0
    st["sig_x"], st["sig_y"], st["sig_z"] = sig
    st["eps_x"], st["eps_y"], st["eps_z"] = eps
    for a, (b, c) in {"ixx": ("sig_y", "sig_z"), "iyy": ("sig_x", "sig_z"), "izz": ("sig_x", "sig_y")}.items():
        st[a] = (st[b] ** 2 + st[c] ** 2) / 20.0
        st[a + "_inv"] = 1.0 / st[a]
    st["target_temp"] = T
    st["dof"] = 6 * n - 3
    return st


def chains(rng, nchains, nbeads, box, bond=1.0):
    """Random-walk chains; returns (positions (n,3), bbi, bbj) 0-based."""
    L = np.asarray(box, dtype=np.float64)
    pos = []
    bi, bj = [], []
    for c in range(nchains):
        p = rng.uniform(-L / 2, L / 2)
        pts = [p]
        for b in range(nbeads - 1):
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            pts.append(pts[-1] + bond * d)
            bi.append(c * nbeads + b)
            bj.append(c * nbeads + b + 1)
        pos += pts
    pos = np.array(pos)
    pos = pos - np.round(pos / L) * L
    return pos, np.array(bi, np.int32), np.array(bj, np.int32)


def to_device(sysdev, st: State, nlist=None):
    """Load an oracle State into a sucrose_amd.System."""
    sysdev.set_num_atoms(st.n, 1)
    sysdev.set_box(st.box[0], st.box[1], st.box[2], periodic=st.periodic)
    for k, f in enumerate(FIELDS):
        sysdev.upload(f, st.buf[:, k])
    for name, idx in [("POT_ENG", "pot_eng"), ("KIN_ENG", "kin_eng"), ("KIN_ROT", "kin_rot"),
                      ("VIRIAL_BONDED", "virial_bonded"), ("TARGET_TEMP", "target_temp"), ("DOF", "dof"),
                      ("DENSITY", "density"), ("HB0", "hb0"), ("HB1", "hb1"), ("HB2", "hb2"),
                      ("ETA0", "eta0"), ("ETA1", "eta1"), ("ETA2", "eta2"),
                      ("PETA0", "peta0"), ("PETA1", "peta1"), ("PETA2", "peta2")]:
        sysdev.set_scalar(name, float(st[idx]))
    if st.nc > 0:
        sysdev.set_constraints(st.bbi + 1, st.bbj + 1, st.bb_len)


def from_device(sysdev, st: State) -> State:
    out = st.copy()
    for k, f in enumerate(FIELDS):
        out.buf[:, k] = sysdev.download(f)
    for name, idx in [("POT_ENG", "pot_eng"), ("KIN_ENG", "kin_eng"), ("KIN_ROT", "kin_rot"),
                      ("VIRIAL_BONDED", "virial_bonded"), ("TEMP", "temp"), ("PRESSURE", "pressure"),
                      ("VIRIAL", "virial"), ("ETA0", "eta0"), ("ETA1", "eta1"), ("ETA2", "eta2"),
                      ("PETA0", "peta0"), ("PETA1", "peta1"), ("PETA2", "peta2"),
                      ("HB0", "hb0"), ("HB1", "hb1"), ("HB2", "hb2")]:
        out[idx] = sysdev.get_scalar(name)
    return out


def rel_err(a, b, scale=None):
    """max |a-b| / scale, scale defaulting to the RMS of b (>0)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if scale is None:
        scale = np.sqrt(np.mean(b * b)) if b.size else 1.0
    scale = scale if scale > 0 else 1.0
    return float(np.max(np.abs(a - b)) / scale) if b.size else 0.0


def report_state(rng, n, L=7.0, nbonds=0):
    """A state for the reporter formats: ordinary values plus the cases the
    F0.d descriptor treats specially (exact ties at 3 and 10 decimals,
    negative values that round to zero, -0.0, |x| < 1, large magnitudes),
    integer types of several widths and signs, and bonds."""
    st = base_state(rng, n, L, sig=(1.0, 1.5, 3.0), eps=(5.0, 4.0, 1.0), charges=True)
    special = np.array([1.0625, -1.0625, 2.0625, -0.0625, 0.0004, -0.0004, -0.0, 0.0, 0.9995, -0.99951,
                        123456.789, -98765.4321, 1e-11, -1e-11, 0.5, -0.5, 2.5e-11, 0.00000000005,
                        -0.00000000005, 1.00000000005])
    fields = ["rx", "ry", "rz", "vx", "vy", "vz", "lx", "ly", "lz", "fx", "fy", "fz", "qx", "charge", "mass"]
    for k, f in enumerate(fields):
        v = st[f].copy()
        m = min(n, len(special))
        v[:m] = np.roll(special, k)[:m] if f != "mass" else np.abs(np.roll(special, k)[:m]) + 0.5
        st[f] = v
    rtype = rng.integers(-3, 2000, n).astype(np.int32)
    rtype[0] = 0
    if nbonds:
        st.bbi = np.arange(nbonds, dtype=np.int64) % (n - 1)
        st.bbj = st.bbi + 1
        st.bb_len = rng.uniform(0.5, 1.5, nbonds)
    st["pot_eng"], st["kin_eng"], st["kin_rot"] = -1234.5625 * n / 8, 0.0625 * n, 0.125 * n
    st["temp"], st["pressure"], st["density"] = 1.0625, -0.00000000005, n / L ** 3
    return st, rtype


def all_pairs(n):
    """A neighbour-list dict holding every pair I < J in the reference's order."""
    return {"offset": np.concatenate([[0], np.cumsum(np.arange(n - 1, -1, -1))]).astype(np.int64),
            "nbr": np.concatenate([np.arange(i + 1, n) for i in range(n)] + [np.zeros(1)]).astype(np.int32),
            "x_last": np.zeros(n), "y_last": np.zeros(n), "z_last": np.zeros(n),
            "cutoff": 0.0, "skin": 0.0, "nnz": n * (n - 1) // 2}


def lattice_lj(rng, n, L, jitter=0.05, T=1.0):
    """n Lennard-Jones particles on a jittered cubic lattice in a box of edge L."""
    m = int(np.ceil(n ** (1 / 3)))
    st = base_state(rng, n, L, forces=False, T=T)
    g = np.arange(m) * (L / m) - L / 2
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)[:n] + rng.normal(size=(n, 3)) * jitter
    st["rx"], st["ry"], st["rz"] = P[:, 0], P[:, 1], P[:, 2]
    return st
