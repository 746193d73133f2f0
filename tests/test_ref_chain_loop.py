"""The reference harness's charged-chain loop (ref_run_chain_steps, the C5
CPU baseline) against the oracle's README loop on the same chains: the
reference's own routines in the loop the benchmark times."""
import numpy as np

from oracle.oracle import DH, LGVC, WCA, Reference
from tests._states import base_state


def test_reference_chain_loop_matches_oracle(oracle, reference):
    import sucrose_amd as sx
    from tests.test_gpu_rows import _bonded_terms, _oracle_bonded
    nchains, nbeads = 10, 20
    n = nchains * nbeads
    L = (n / 0.5) ** (1 / 3)
    rng = np.random.default_rng(3)
    st = base_state(rng, n, L, forces=False, T=1.0, charges=True)
    st["charge"] = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    pos = sx.gen_chains(nchains, nbeads, L, bond=1.0, min_dist=0.95, seed=5)
    st["rx"], st["ry"], st["rz"] = pos[:, 0], pos[:, 1], pos[:, 2]
    ids = np.arange(n).reshape(nchains, nbeads)
    st.bbi = ids[:, :-1].ravel().astype(np.int32)
    st.bbj = ids[:, 1:].ravel().astype(np.int32)
    st.bb_len = np.ones(len(st.bbi))
    st["dof"] = 6 * n - st.nc - 3
    terms = _bonded_terms(nchains, nbeads)
    R = Reference(private=True)
    R.load(st)
    R.nlist_init(0.3, 3.0)
    R.run_chain_steps(5, 1, 0.005, 1.0, (1.0, 1.0), (3.0, 1.0, 1.0), terms[0][1], 20.0, 180.0, terms[1][1], 1.0)
    got = st.copy()
    R.store(got)
    ref = st.copy()
    nl = oracle.nlist_new(ref, 3.0, 0.3)
    oracle.nlist_build(ref, nl)
    for t in range(5):
        oracle.integrate(LGVC, 0, ref, 0.005, step=1 + t, damping=1.0)
        if oracle.needs_rebuild(ref, nl):
            oracle.nlist_build(ref, nl)
        oracle.call("or_initialize_forces", ref)
        oracle.pair_loop(ref, nl, WCA, [1.0, 1.0])
        oracle.pair_loop(ref, nl, DH, [3.0, 1.0, 1.0])
        _oracle_bonded(oracle, ref, terms)
        oracle.integrate(LGVC, 1, ref, 0.005, damping=1.0)
        oracle.call("or_calculate_thermo", ref)
    # one pass calling both pair routines vs two loops: summation order only
    for f in ("rx", "ry", "vx", "vz", "fx", "fy"):
        e = np.max(np.abs(got[f] - ref[f])) / np.sqrt(np.mean(ref[f] ** 2))
        assert e < 1e-12, (f, e)
    assert abs(got["pot_eng"] - ref["pot_eng"]) <= 1e-12 * abs(ref["pot_eng"])
