"""Reference-vs-oracle jobs that need a fresh process each.

The reference keeps its neighbour list in module globals that only
``init_nlist`` allocates (fortran/system.f90:895-964) and a second call in the
same process fails, so every job that touches the list runs in a spawned
child.  Each job returns plain Python data.
"""
from __future__ import annotations

import concurrent.futures as cf
import multiprocessing as mp

import numpy as np


def in_child(fn, *args):
    with cf.ProcessPoolExecutor(max_workers=1, mp_context=mp.get_context("spawn")) as ex:
        return ex.submit(fn, *args).result(timeout=600)


def _positions(st, rng, box):
    for k, f in enumerate(["rx", "ry", "rz"]):
        st[f] = rng.uniform(-box[k] / 2, box[k] / 2, st.n)


def nlist_job(seed, n, box, rc, skin, moves):
    """Build the list on both sides, then move particles `moves` times and
    compare the rebuild decision and the rebuilt list after each move."""
    import ctypes as C

    from oracle.oracle import Oracle, Reference, State
    O, R = Oracle(), Reference()
    rng = np.random.default_rng(seed)
    st = State(n, box)
    _positions(st, rng, st.box)
    R.load(st)
    R.nlist_init(skin, rc)
    nl = O.nlist_new(st, rc, skin)
    O.nlist_build(st, nl)
    out = []

    def compare():
        off, nbr, nnz = R.nlist_get(cap=max(64, 60 * n))
        same = (nnz == nl["nnz"] and np.array_equal(nbr - 1, nl["nbr"][:nnz])
                and np.array_equal(off[:-1] - 1, nl["offset"][:n - 1]) and off[-1] - 1 == nnz)
        return bool(same), int(nnz)

    out.append(("build",) + compare())
    for t, scale in enumerate(moves):
        for f in ["rx", "ry", "rz"]:
            st[f] = st[f] + rng.normal(size=n) * scale
        buf = np.asfortranarray(st.buf)
        R.lib.ref_set_state(C.c_int(n), buf.ctypes.data_as(C.POINTER(C.c_double)))
        a, b = R.needs_rebuild(), O.needs_rebuild(st, nl)
        if a:
            R.nlist_build()
        if b:
            O.nlist_build(st, nl)
        out.append((f"move{t}", a == b) + compare() + (bool(a),))
    return out


def run_job(integ, kind, p, n, box, seed, nsteps, dt, damping, sig, eps, rc, skin, T):
    """The whole README loop on both sides; returns (bit-equal, max |diff|)."""
    from oracle.oracle import Oracle, Reference
    from tests._loop import oracle_run_steps
    from tests._states import base_state
    O, R = Oracle(), Reference()
    rng = np.random.default_rng(seed)
    st = base_state(rng, n, box, sig=sig, eps=eps, T=T, forces=False)
    # spread particles on a jittered lattice so short runs stay finite
    m = int(np.ceil(n ** (1 / 3)))
    g = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
    pos = (g + 0.5) / m * box - box / 2 + rng.normal(size=(n, 3)) * 0.02
    st["rx"], st["ry"], st["rz"] = pos[:, 0], pos[:, 1], pos[:, 2]
    a = st.copy()
    R.load(st)
    R.nlist_init(skin, rc)
    R.run_steps(integ, kind, p, nsteps, 1, dt, damping)
    b = st.copy()
    R.store(b)
    nl = O.nlist_new(a, rc, skin)
    O.nlist_build(a, nl)
    oracle_run_steps(O, a, nl, integ, kind, p, nsteps, 1, dt, damping)
    same = bool(np.array_equal(a.buf, b.buf) and np.array_equal(a.scal, b.scal))
    diff = float(max(np.abs(a.buf - b.buf).max(), np.abs(a.scal - b.scal).max()))
    return same, diff, float(b["pot_eng"]), float(b["temp"])
