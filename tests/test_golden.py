"""The oracle against the reference's own outputs stored in
tests/golden/reference_vectors.npz (written by tests/golden/make_golden.py
from the reference compiled from source).  Runs without the reference build.

Integer-valued and purely arithmetic results are compared exactly; results
that pass through libm (exp, log, sqrt, sin, cos, pow) to 1e-13 relative, the
allowance for a different libm build on another machine.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle.oracle import (DH, ECP, ECPR, GB, GBR, LGV, LGVC, LJ, LJ12, LJW, MORSE, NHT, QQ, VV, VVC,
                           WCA, State)

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))
PD = C.POINTER(C.c_double)
TOL = 1e-13


def close(a, b, tol=TOL):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(np.abs(b).max(), 1e-300) if b.size else 1.0
    return a.shape == b.shape and (np.abs(a - b).max() <= tol * scale if b.size else True)


def test_philox(oracle):
    for c, u, nrm in zip(G["philox_ctr"], G["philox_uniform"], G["philox_normal"]):
        out = np.zeros(4)
        oracle.lib.or_philox_uniform((C.c_int32 * 4)(*[int(x) for x in c]), out.ctypes.data_as(PD))
        assert np.array_equal(out, u)
        assert close(oracle.philox_normal(*[int(x) for x in c]), nrm)


def test_brent(oracle):
    for ar, br, r, lam in zip(G["brent_ar"], G["brent_br"], G["brent_r"], G["brent_lambda"]):
        ar, br, r = (np.ascontiguousarray(x) for x in (ar, br, r))
        nit = C.c_int()
        got = oracle.lib.or_ecp_brent(ar.ctypes.data_as(PD), br.ctypes.data_as(PD), r.ctypes.data_as(PD),
                                      C.byref(nit))
        assert abs(got - lam) <= 1e-13


KINDS = {"lj": LJ, "lj12": LJ12, "ljw": LJW, "wca": WCA, "morse": MORSE, "dh": DH, "gb": GB, "gbr": GBR,
         "ecp": ECP, "ecpr": ECPR}


@pytest.mark.parametrize("name", list(KINDS))
def test_pairs(oracle, name):
    buf = G[f"pair_{name}_buf"]
    st = State(buf.shape[0], G[f"pair_{name}_box"], buf=np.asfortranarray(buf.copy()))
    p = list(G[f"pair_{name}_p"])
    out = G[f"pair_{name}_out"]
    assert np.count_nonzero(out[:, 0]) > 8
    for k in range(out.shape[0]):
        r = oracle.pair(KINDS[name], st, 2 * k, 2 * k + 1, p)
        got = np.concatenate([[r[0]], r[1], r[2], r[3]])
        assert close(got, out[k]), (name, k)


@pytest.mark.parametrize("name,kind,coupling", [("vv", VV, None), ("qq", QQ, None), ("nht", NHT, None),
                                                ("nht_c", NHT, 2.5), ("lgv", LGV, None)])
@pytest.mark.parametrize("phase", [0, 1])
def test_integrators(oracle, name, kind, coupling, phase):
    k = f"int_{name}{phase}"
    st = State(G[k + "_in_buf"].shape[0], G[k + "_box"], buf=np.asfortranarray(G[k + "_in_buf"].copy()),
               scal=G[k + "_in_scal"].copy())
    oracle.integrate(kind, phase, st, 0.003, step=7, damping=1.3, coupling=coupling)
    assert close(st.buf, G[k + "_out_buf"]) and close(st.scal, G[k + "_out_scal"])


@pytest.mark.parametrize("name,kind", [("vvc", VVC), ("lgvc", LGVC)])
@pytest.mark.parametrize("phase", [0, 1])
def test_constrained(oracle, name, kind, phase):
    bi, bj = G["chain_bbi"], G["chain_bbj"]
    st = State(G["chain_in_buf"].shape[0], G["chain_box"], buf=np.asfortranarray(G["chain_in_buf"].copy()),
               scal=G["chain_in_scal"].copy(), bbi=bi.copy(), bbj=bj.copy(), bb_len=np.ones(len(bi)))
    oracle.integrate(kind, phase, st, 0.005, step=3, damping=1.0)
    assert close(st.buf, G[f"chain_{name}{phase}_buf"]) and close(st.scal, G[f"chain_{name}{phase}_scal"])


def test_nlist(oracle):
    L, rc, skin = G["nlist_params"]
    pos = G["nlist_pos"]
    st = State(pos.shape[0], [L, L, L])
    st["rx"], st["ry"], st["rz"] = pos[:, 0], pos[:, 1], pos[:, 2]
    nl = oracle.nlist_build(st, oracle.nlist_new(st, rc, skin))
    off1, nbr1 = G["nlist_offset1"], G["nlist_nbr1"]
    nnz = nl["nnz"]
    assert nnz == len(nbr1) == off1[-1] - 1
    assert np.array_equal(nl["nbr"][:nnz] + 1, nbr1)
    assert np.array_equal(nl["offset"][:st.n - 1] + 1, off1[:-1])
