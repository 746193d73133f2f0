"""The minimum-image helper of the device kernels (sx_mimg in
sucrose_amd/csrc/sx_internal.h) returns d + 0.0 without dividing when
|d| < L/2, in place of d - anint(d/L)*L.  The two agree bit for bit exactly
when the correctly rounded quotient d/L stays below 1/2 in magnitude there
(anint then gives a signed zero, and d - (+-0)*L equals d + 0.0, -0.0
included).  IEEE division is correctly rounded in numpy as on the device, so
the quotient property is checked here on random boxes and on the hardest
inputs: the doubles just below L/2 and just above -L/2."""
import numpy as np


def _anint(x):
    """C round(): half away from zero, signed zeros kept; exact (x - trunc(x)
    is exact, unlike floor(x + 0.5), which rounds 0.5 - 2**-54 up to 1)."""
    t = np.trunc(x)
    return np.where(np.abs(x - t) >= 0.5, t + np.copysign(1.0, x), t)


def _full(d, L):
    return d - _anint(d / L) * L


def test_quotient_below_half_inside_half_box():
    rng = np.random.default_rng(1)
    L = np.concatenate([rng.uniform(0.5, 500.0, 200000), 2.0 ** rng.integers(-20, 20, 20000).astype(float),
                        np.nextafter(2.0 ** rng.integers(-20, 20, 20000).astype(float), 0.0)])
    half = 0.5 * L
    edge = np.nextafter(half, 0.0)                    # largest |d| below L/2
    cases = [edge, -edge, np.nextafter(edge, 0.0), rng.uniform(-1, 1, L.size) * half]
    for d in cases:
        inside = np.abs(d) < half
        assert inside.all()
        q = d / L
        assert np.all(np.abs(q) < 0.5)
        assert np.all(_anint(q) == 0.0)
        got = d + 0.0
        want = _full(d, L)
        assert np.array_equal(got.view(np.int64), want.view(np.int64))


def test_negative_zero():
    d = np.array([-0.0, 0.0])
    L = np.array([3.0, 3.0])
    want = _full(d, L)
    got = d + 0.0
    assert np.array_equal(np.signbit(got), np.signbit(want)) and not np.signbit(got).any()
