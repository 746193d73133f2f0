"""The allocation-free entry point sx_philox_normals_noalloc against the entry
point it varies, sx_philox_normals, both called through the C ABI: two calls
in a row, the second with fewer normals than the first (the scratch then
still holds the first call's values past the second's), then the size limit.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(L, h, name, n, c2, c3, c4):
    out = np.full(3 * n, np.nan)
    rc = getattr(L, name)(h, n, c2, c3, c4, out.ctypes.data_as(C.POINTER(C.c_double)))
    return rc, out


def test_philox_normals_noalloc_matches_allocating_call():
    import sucrose_amd as sx
    L = sx.lib()
    h = C.c_void_p()
    assert L.sx_create(0, C.byref(h)) == 0
    try:
        for n, (c2, c3, c4) in ((150000, (17, 1500, 1)), (4097, (3, 1000, 0))):
            print(f"philox normals n={n} counters ({c2}, {c3}, {c4})", flush=True)
            rc_a, a = _call(L, h, "sx_philox_normals", n, c2, c3, c4)
            rc_b, b = _call(L, h, "sx_philox_normals_noalloc", n, c2, c3, c4)
            assert rc_a == 0 and rc_b == 0, (rc_a, rc_b)
            assert np.array_equal(a, b)
        # the largest request that fits the 4 MiB scratch, then one beyond it:
        # a clear error and nothing written
        rc_a, a = _call(L, h, "sx_philox_normals", 174762, 5, 2000, 1)
        rc_b, b = _call(L, h, "sx_philox_normals_noalloc", 174762, 5, 2000, 1)
        assert rc_a == 0 and rc_b == 0 and np.array_equal(a, b)
        rc, out = _call(L, h, "sx_philox_normals_noalloc", 174763, 5, 2000, 1)
        assert rc == -4 and np.all(np.isnan(out))
        assert b"174762" in L.sx_last_error(h)
    finally:
        assert L.sx_destroy(h) == 0
