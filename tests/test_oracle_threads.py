"""The all-cores variants of the oracle's pair loop and list build (the
threaded CPU baseline) give the serial oracle's results bit for bit."""
import os

import numpy as np
import pytest

from oracle.oracle import ECPR, GB, LJ, Oracle
from tests._states import base_state

OMP = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "liboracle_omp.so")


@pytest.mark.skipif(not os.path.exists(OMP), reason="liboracle_omp.so not built")
@pytest.mark.parametrize("kind,params,kw,rc", [(LJ, [1.0, 1.0, 2.5], {}, 2.5),
                                               (GB, [4.0, 1, 2], dict(sig=(1, 1, 3), eps=(5, 5, 1)), 6.0),
                                               (ECPR, [0.0], dict(sig=(1, 1, 3)), 3.1)])
def test_threaded_loop_equals_serial(oracle, kind, params, kw, rc):
    T = Oracle(OMP)
    rng = np.random.default_rng(5 + kind)
    st = base_state(rng, 700, 11.0, forces=False, **kw)
    a, b = st.copy(), st.copy()
    nla = oracle.nlist_new(a, rc, 0.3)
    oracle.nlist_build(a, nla)
    nlb = T.nlist_new(b, rc, 0.3)
    T.nlist_build_threads(b, nlb)
    assert nla["nnz"] == nlb["nnz"] > 0
    assert np.array_equal(nla["offset"], nlb["offset"])
    assert np.array_equal(nla["nbr"][:nla["nnz"]], nlb["nbr"][:nlb["nnz"]])
    oracle.pair_loop(a, nla, kind, params)
    T.pair_loop_threads(b, nlb, kind, params)
    assert np.array_equal(a.buf, b.buf)
    assert a["pot_eng"] == b["pot_eng"]
