"""The C-ABI library: it loads, exports every entry point include/sucrose.h
declares, and its host-only routines agree with the oracle.  No device
compute is called here (no GPU is needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sucrose.h")
LIB = os.path.join(ROOT, "sucrose_amd", "libsucrose_hip.so")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built (__graft_entry__.build())")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|void|long|double|char)\s*\*?\s*(sx_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared()
    for must in ["sx_create", "sx_destroy", "sx_upload", "sx_download", "sx_nlist_init", "sx_nlist_build",
                 "sx_nlist_needs_rebuild", "sx_forces", "sx_pair_eval", "sx_vv_initial_step",
                 "sx_vv_final_step", "sx_qq_initial_step", "sx_qq_final_step", "sx_nht_initial_step",
                 "sx_nht_final_step", "sx_lgv_initial_step", "sx_lgv_final_step", "sx_apply_constraints_a",
                 "sx_apply_constraints_b", "sx_calculate_thermo", "sx_run", "sx_philox4x32_10"]:
        assert must in names
    assert len(names) > 50


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_no_device_error():
    import sucrose_amd as sx
    L = sx.lib()
    assert L.sx_version() > 0
    # without a usable GPU, sx_create reports an error instead of aborting
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    p = C.c_void_p()
    assert L.sx_create(0, C.byref(p)) < 0
    assert L.sx_last_error(None)


def test_host_philox_matches_oracle(oracle):
    import sucrose_amd as sx
    rng = np.random.default_rng(0)
    for ctr in rng.integers(0, 2**32, size=(300, 4), dtype=np.uint64):
        assert np.array_equal(sx.philox4x32_10([int(x) for x in ctr]), oracle.philox_words(ctr))


def test_host_workload_generator(oracle):
    """The random-sequential-addition generator places non-overlapping
    ellipsoids: for every pair closer than the longest axis the
    Perram-Wertheim contact function max_lambda F(lambda) of the oracle is
    at least 1 (F < 1 is overlap), and the library's own counter agrees."""
    from scipy.optimize import minimize_scalar

    import sucrose_amd as sx
    from oracle.oracle import State
    n, L = 300, 14.0
    pos, quat = sx.gen_ellipsoids_rsa(n, (0.5, 0.5, 1.5), L, seed=3)
    assert pos.shape == (n, 3) and quat.shape == (n, 4)
    assert np.all(np.abs(pos) <= L / 2 + 1e-12)
    assert np.allclose(np.linalg.norm(quat, axis=1), 1.0)
    st = State(n, [L, L, L])
    for k, f in enumerate(["rx", "ry", "rz"]):
        st[f] = pos[:, k]
    for k, f in enumerate(["qw", "qx", "qy", "qz"]):
        st[f] = quat[:, k]
    st["sig_x"], st["sig_y"], st["sig_z"] = 1.0, 1.0, 3.0
    d = pos[:, None, :] - pos[None, :, :]
    d -= np.round(d / L) * L
    close_i, close_j = np.nonzero(np.triu((d ** 2).sum(-1) < 9.0, 1))
    assert len(close_i) > 50
    for i, j in zip(close_i, close_j):
        ar, br, r = oracle.ecp_shape(st, i, j)
        res = minimize_scalar(lambda lam: oracle.optim_dist(lam, ar, br, r), bounds=(0.0, 1.0),
                              method="bounded", options={"xatol": 1e-10})
        assert -res.fun >= 1.0 - 1e-9, (i, j, -res.fun)
    assert sx.count_ellipsoid_overlaps(pos, quat, (0.5, 0.5, 1.5), L) == 0
