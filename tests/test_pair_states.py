"""The per-pair parity cases cover what they claim (CPU, oracle only).

The device tests in test_gpu_parity.py compare every pair kind on states
built by tests/_pairs.py.  This checks those states: most pairs interact,
each switch band and weld band holds pairs, and no pair is deeply overlapped
(the largest |force| is within 1e4 times the median; random placement gave 1e13 and more), so that the
element-wise tolerance of the device tests applies to ordinary pairs.
"""
import numpy as np
import pytest

from oracle.oracle import COULOMB, ECPR, GBR
from tests._pairs import ECP_KINDS, band_counts, pair_state
from tests.test_gpu_parity import PAIR_CASES
from tests._states import base_state


@pytest.mark.parametrize("kind,params,kw", PAIR_CASES)
def test_pair_cases_cover_bands(oracle, kind, params, kw):
    rng = np.random.default_rng(100 + kind)
    npairs = 400
    st, i, j, d, contact = pair_state(oracle, rng, kind, params, npairs,
                                      lambda r, n, L: base_state(r, n, L, **kw))
    p = list(params) if kind not in (GBR, ECPR) else [0.0] + list(params[1:])
    pe = np.zeros(npairs)
    fm = np.zeros(npairs)
    for t in range(npairs):
        r = oracle.pair(kind, st, i[t], j[t], p if kind != ECPR else [0.0])
        pe[t] = r[0]
        fm[t] = np.linalg.norm(r[1])
    nz = np.count_nonzero(pe)
    assert nz >= npairs // 3, nz
    med = np.median(fm[fm > 0])
    assert fm.max() <= 1e4 * med, (fm.max(), med)
    bands = band_counts(kind, params, d, contact)
    for name, cnt in bands.items():
        assert cnt >= npairs // 20, (name, cnt)
    if kind == COULOMB:
        assert bands["switch"] > 0
