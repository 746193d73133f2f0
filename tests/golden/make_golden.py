"""Writes tests/golden/reference_vectors.npz from the SUCROSE reference.

The reference is compiled from its own Fortran sources by ``make -C oracle
ref`` (flang, -fdefault-real-8, IEEE order); this script drives it through
oracle/ref_harness.f90 on small seeded inputs and stores inputs and outputs.
The fixture lets the oracle be checked where the reference build is absent.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import (DH, ECP, ECPR, GB, GBR, LGV, LGVC, LJ, LJ12, LJW, MORSE, NHT,  # noqa: E402
                           QQ, VV, VVC, WCA, Reference, State)
from tests._refjobs import in_child  # noqa: E402
from tests._states import base_state, chains, rand_quats  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_vectors.npz")

PAIR_SPECS = [
    ("lj", LJ, [1.0, 1.0, 2.5], dict(), 0.8, 2.8),
    ("lj12", LJ12, [1.0, 1.0, 2.5], dict(), 0.8, 2.8),
    ("ljw", LJW, [1.0, 1.0, 2.5, 0.3], dict(), 0.8, 2.8),
    ("wca", WCA, [1.0, 1.0], dict(), 0.8, 1.3),
    ("morse", MORSE, [1.0, 2.0, 1.1, 2.5], dict(), 0.6, 2.8),
    ("dh", DH, [5.0, 1.0, 1.0], dict(charges=True), 0.5, 5.5),
    ("gb", GB, [4.0, 1, 2], dict(sig=(1, 1, 3), eps=(5, 5, 1)), 0.9, 4.5),
    ("gbr", GBR, [0.0, 1, 2], dict(sig=(1, 1, 3), eps=(5, 5, 1)), 0.9, 2.5),
    ("ecp", ECP, [1.5], dict(sig=(1, 1, 3)), 0.8, 3.4),
    ("ecpr", ECPR, [0.0], dict(sig=(1, 1, 3)), 0.8, 3.4),
]


def pair_state(rng, npairs, kw, rmin, rmax, box=40.0):
    st = base_state(rng, 2 * npairs, box, **kw)
    d = rng.normal(size=(npairs, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = rng.uniform(rmin, rmax, npairs)
    c = rng.uniform(-box / 2, box / 2, (npairs, 3))
    c[: npairs // 4, 0] = box / 2 - 0.1
    p2 = c + d * r[:, None]
    p2 -= np.round(p2 / box) * box
    for k, f in enumerate(["rx", "ry", "rz"]):
        st[f][0::2] = c[:, k]
        st[f][1::2] = p2[:, k]
    return st


def nlist_vectors(seed, n, L, rc, skin):
    R = Reference()
    rng = np.random.default_rng(seed)
    st = State(n, [L, L, L])
    for f in ["rx", "ry", "rz"]:
        st[f] = rng.uniform(-L / 2, L / 2, n)
    R.load(st)
    R.nlist_init(skin, rc)
    off, nbr, nnz = R.nlist_get(cap=64 * n)
    return st.buf[:, :3].copy(), off, nbr


def main():
    R = Reference()
    g = {}
    rng = np.random.default_rng(2024)
    # Philox uniforms and Box-Muller normals (xmath.f90:40-111)
    ctr = rng.integers(-2**31, 2**31, size=(512, 4)).astype(np.int64)
    g["philox_ctr"] = ctr
    g["philox_uniform"] = np.array([R.philox_uniform(c) for c in ctr])
    g["philox_normal"] = np.array([R.philox_normal(*[int(x) for x in c]) for c in ctr])
    # Brent contact parameter (force.f90:1153-1200)
    ars, brs, rs, lams = [], [], [], []
    for _ in range(128):
        q = rand_quats(rng, 2)
        mats = []
        for qq in q:
            w, x, y, z = qq
            A = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)],
                          [2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)],
                          [2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)]])
            mats.append(np.asfortranarray(A.T @ np.diag(rng.uniform(0.25, 2.0, 3) ** 2) @ A))
        r = rng.normal(size=3) * rng.uniform(0.5, 4.0)
        ars.append(mats[0].ravel(order="F")); brs.append(mats[1].ravel(order="F")); rs.append(r)
        lams.append(R.brent(mats[0], mats[1], r))
    g["brent_ar"], g["brent_br"], g["brent_r"], g["brent_lambda"] = map(np.array, (ars, brs, rs, lams))
    # pair kernels
    for name, kind, p, kw, rmin, rmax in PAIR_SPECS:
        st = pair_state(rng, 64, kw, rmin, rmax)
        R.load(st)
        outs = np.array([np.concatenate([[a[0]], a[1], a[2], a[3]])
                         for a in (R.pair(kind, 2 * k, 2 * k + 1, p) for k in range(64))])
        g[f"pair_{name}_buf"] = st.buf.copy()
        g[f"pair_{name}_box"] = st.box.copy()
        g[f"pair_{name}_p"] = np.array(p, dtype=np.float64)
        g[f"pair_{name}_out"] = outs           # pe, f(3), ti(3), tj(3)
    # one phase of each integrator on 32 particles
    for name, kind, coupling in [("vv", VV, None), ("qq", QQ, None), ("nht", NHT, None),
                                 ("nht_c", NHT, 2.5), ("lgv", LGV, None)]:
        for phase in (0, 1):
            st = base_state(rng, 32, 5.0, sig=(1, 1, 3), eps=(5, 5, 1), T=1.5)
            st["peta0"], st["peta1"], st["peta2"] = 0.3, -0.2, 0.1
            st["eta0"], st["eta1"], st["eta2"] = 0.05, 0.01, -0.02
            st["kin_eng"], st["kin_rot"] = 24.0, 20.0
            R.load(st)
            R.integrate(kind, phase, 0.003, step=7, damping=1.3, coupling=coupling)
            out = st.copy()
            R.store(out)
            g[f"int_{name}{phase}_in_buf"], g[f"int_{name}{phase}_in_scal"] = st.buf.copy(), st.scal.copy()
            g[f"int_{name}{phase}_out_buf"], g[f"int_{name}{phase}_out_scal"] = out.buf.copy(), out.scal.copy()
            g[f"int_{name}{phase}_box"] = st.box.copy()
    # constrained steps and RATTLE on 4 chains of 8 beads
    L = 7.0
    pos, bi, bj = chains(rng, 4, 8, [L, L, L])
    st = base_state(rng, len(pos), L, T=1.0)
    st["rx"], st["ry"], st["rz"] = pos[:, 0], pos[:, 1], pos[:, 2]
    st.bbi, st.bbj, st.bb_len = bi, bj, np.ones(len(bi))
    st["dof"] = 6 * st.n - len(bi) - 3
    g["chain_bbi"], g["chain_bbj"], g["chain_box"] = bi, bj, st.box.copy()
    g["chain_in_buf"], g["chain_in_scal"] = st.buf.copy(), st.scal.copy()
    for name, kind in [("vvc", VVC), ("lgvc", LGVC)]:
        for phase in (0, 1):
            R.load(st)
            R.integrate(kind, phase, 0.005, step=3, damping=1.0)
            out = st.copy()
            R.store(out)
            g[f"chain_{name}{phase}_buf"], g[f"chain_{name}{phase}_scal"] = out.buf.copy(), out.scal.copy()
    # neighbour list CSR, 256 particles (fresh process: init_nlist runs once)
    pos, off, nbr = in_child(nlist_vectors, 7, 256, 8.0, 2.5, 0.3)
    g["nlist_pos"], g["nlist_offset1"], g["nlist_nbr1"] = pos, off, nbr
    g["nlist_params"] = np.array([8.0, 2.5, 0.3])
    np.savez_compressed(OUT, **g)
    print(f"wrote {OUT}: {os.path.getsize(OUT) / 1024:.0f} KiB, {len(g)} arrays")


if __name__ == "__main__":
    main()
