"""Pins of the C oracle that do not need the reference build.

Each test checks one restated routine against something the mathematics or a
published source fixes independently of the restatement: Random123's
known-answer vectors for Philox4x32-10, numpy's linear algebra, scipy's Brent
minimiser run with the reference's constants, brute-force neighbour search,
central finite differences of the pair energies (forces are -dU/dr_i and
torques -dU/dtheta for a rotation of particle i about a space axis), closed
forms for spheres, and conservation laws of the integrators.
"""
import ctypes as C

import numpy as np
import pytest

from oracle.oracle import (COULOMB, DH, ECP, ECPC, ECPR, ECPW, ECPWC, GB, GBC, GBR, GBW, GBWC, LGV, LJ, LJ12, LJW, MORSE, QQ,
                           VV, VVC, WCA, State)
from tests._states import base_state, chains, rand_quats

PD = C.POINTER(C.c_double)
REF_KEY = (0x6D2B79F5, 0x1B873593)   # xmath.f90:10-11


def test_philox_random123_kats(oracle):
    # Random123 kat_vectors, philox4x32 10 rounds: (ctr, key) -> out
    kats = [
        ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
        ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
        ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
         (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
    ]
    for ctr, key, out in kats:
        assert list(oracle.philox_words_key(ctr, key)) == list(out)


def test_philox_reference_key(oracle):
    rng = np.random.default_rng(0)
    for ctr in rng.integers(0, 2**32, size=(200, 4), dtype=np.uint64):
        assert np.array_equal(oracle.philox_words(ctr), oracle.philox_words_key(ctr, REF_KEY))


def test_philox_uniform_is_word_over_2_32(oracle):
    rng = np.random.default_rng(1)
    for ctr in rng.integers(-2**31, 2**31, size=(200, 4)):
        u = np.zeros(4)
        oracle.lib.or_philox_uniform((C.c_int32 * 4)(*[int(x) for x in ctr]), u.ctypes.data_as(PD))
        w = oracle.philox_words([int(x) & 0xFFFFFFFF for x in ctr])
        assert np.array_equal(u, w.astype(np.float64) / 4294967296.0)


def test_box_muller_radius_and_moments(oracle):
    """n1, n2 come from (u1, u2); n1^2 + n2^2 = -2 log u1 whatever the angle."""
    x = np.array([oracle.philox_normal(i, 3, 5, 7) for i in range(1, 20001)])
    for i in range(1, 50):
        w = oracle.philox_words([i, 3, 5, 7]).astype(np.float64) / 4294967296.0
        assert abs(x[i - 1, 0] ** 2 + x[i - 1, 1] ** 2 + 2 * np.log(max(w[0], 1e-35))) < 1e-12
    assert np.all(np.abs(x.mean(0)) < 0.03)
    assert np.all(np.abs(x.var(0) - 1) < 0.04)
    assert abs(np.corrcoef(x[:, 0], x[:, 1])[0, 1]) < 0.03


def test_linear_algebra(oracle):
    rng = np.random.default_rng(2)
    for _ in range(200):
        m = rng.normal(size=(3, 3))
        a = m @ m.T + 0.5 * np.eye(3)
        b = rng.normal(size=3)
        o = np.zeros(3)
        af = np.asfortranarray(a)
        oracle.lib.or_cholesky_solve(af.ctypes.data_as(PD), b.ctypes.data_as(PD), o.ctypes.data_as(PD))
        assert np.allclose(o, np.linalg.solve(a, b), rtol=1e-10, atol=1e-12)
        g = np.asfortranarray(rng.normal(size=(3, 3)))
        inv = np.zeros(9)
        oracle.lib.or_inverse(g.ctypes.data_as(PD), inv.ctypes.data_as(PD))
        assert np.allclose(inv.reshape(3, 3, order="F"), np.linalg.inv(g), rtol=1e-9, atol=1e-9)
        oracle.lib.or_determinant.restype = C.c_double
        assert np.isclose(oracle.lib.or_determinant(g.ctypes.data_as(PD)), np.linalg.det(g), rtol=1e-12)


def _shape_pairs(rng, count):
    out = []
    for _ in range(count):
        q = rand_quats(rng, 2)
        mats = []
        for qq in q:
            w, x, y, z = qq
            A = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)],
                          [2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)],
                          [2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)]])
            mats.append(np.asfortranarray(A.T @ np.diag(rng.uniform(0.25, 2.0, 3) ** 2) @ A))
        out.append((mats[0], mats[1], rng.normal(size=3) * rng.uniform(0.5, 4.0)))
    return out


def test_ecp_brent_is_scipy_brent(oracle):
    """The reference's Brent (force.f90:1153-1200) follows scipy's Brent
    minimiser step for step; with tol=1e-2, _mintol=1e-3, the exact golden
    ratio constant and the bracket (0, psi, 1) they agree to the last bit."""
    from scipy.optimize._optimize import Brent
    psi = 0.5 * (3 - 5 ** 0.5)
    for ar, br, r in _shape_pairs(np.random.default_rng(3), 300):
        f = lambda lam: oracle.optim_dist(lam, ar, br, r)  # noqa: E731
        b = Brent(f, tol=1e-2, maxiter=10000)
        b._mintol = 1e-3
        b._cg = psi
        b.set_bracket((0.0, psi, 1.0))
        b.optimize()
        nit = C.c_int()
        lam = oracle.lib.or_ecp_brent(ar.ctypes.data_as(PD), br.ctypes.data_as(PD), r.ctypes.data_as(PD),
                                      C.byref(nit))
        assert lam == b.xmin and nit.value == b.iter


def test_contact_function_spheres(oracle):
    """Spheres of radii a, b: -F(lam) = lam(1-lam) r^2 / ((1-lam) a^2 + lam b^2),
    maximised at lam* = a/(a+b) with F(lam*) = r^2/(a+b)^2."""
    rng = np.random.default_rng(4)
    for _ in range(100):
        a, b = rng.uniform(0.3, 2.0, 2)
        ar = np.asfortranarray(np.eye(3) * a * a)
        br = np.asfortranarray(np.eye(3) * b * b)
        r = rng.normal(size=3) * 3
        r2 = r @ r
        for lam in (0.1, 0.37, 0.5, 0.9):
            val = -oracle.optim_dist(lam, ar.ravel(order="F"), br.ravel(order="F"), r)
            assert np.isclose(val, lam * (1 - lam) * r2 / ((1 - lam) * a * a + lam * b * b), rtol=1e-12)
        ls = a / (a + b)
        assert np.isclose(-oracle.optim_dist(ls, ar.ravel(order="F"), br.ravel(order="F"), r),
                          r2 / (a + b) ** 2, rtol=1e-12)


def _brute_pairs(st, rlist):
    pos = np.stack([st["rx"], st["ry"], st["rz"]], 1)
    d = pos[None, :, :] - pos[:, None, :]
    d -= np.round(d / st.box) * st.box
    r2 = (d ** 2).sum(-1)
    i, j = np.nonzero(np.triu(r2 < rlist * rlist, 1))
    return list(zip(i.tolist(), j.tolist()))


@pytest.mark.parametrize("n,box,rc,skin", [(300, [9.0, 9.0, 9.0], 2.5, 0.3), (400, [7.0, 11.0, 8.0], 2.0, 0.4)])
def test_nlist_is_brute_force(oracle, n, box, rc, skin):
    """build_nlist (system.f90:966-1003): every pair i<j closer than rc+skin
    under the minimum image, j ascending within each row."""
    rng = np.random.default_rng(5)
    st = State(n, box)
    for k, f in enumerate(["rx", "ry", "rz"]):
        st[f] = rng.uniform(-box[k] / 2, box[k] / 2, n)
    nl = oracle.nlist_build(st, oracle.nlist_new(st, rc, skin))
    off = nl["offset"]
    got = [(i, int(nl["nbr"][k])) for i in range(n) for k in range(off[i], off[i + 1])]
    assert got == _brute_pairs(st, rc + skin)


def test_rebuild_threshold(oracle):
    """needs_rebuild (system.f90:1005-1036) fires when 4 max|dr|^2 exceeds
    (rc+skin)^2 -- a displacement of half the list radius, not half the skin."""
    st = State(50, [20.0] * 3)
    st["rx"] = np.linspace(-9, 9, 50)
    nl = oracle.nlist_build(st, oracle.nlist_new(st, 2.5, 0.3))
    half = (2.5 + 0.3) / 2
    for d, want in [(0.2, False), (half * (1 - 1e-9), False), (half * (1 + 1e-9), True)]:
        s2 = st.copy()
        s2["ry"][7] += d
        assert oracle.needs_rebuild(s2, nl) == want, d


def _rotated(st, i, axis, angle):
    w, x, y, z = st["qw"][i], st["qx"][i], st["qy"][i], st["qz"][i]
    dq = np.zeros(4)
    dq[0] = np.cos(angle / 2)
    dq[1 + axis] = np.sin(angle / 2)
    a0, a1, a2, a3 = dq
    q = [a0 * w - a1 * x - a2 * y - a3 * z, a0 * x + a1 * w + a2 * z - a3 * y,
         a0 * y - a1 * z + a2 * w + a3 * x, a0 * z + a1 * y - a2 * x + a3 * w]
    s2 = st.copy()
    s2["qw"][i], s2["qx"][i], s2["qy"][i], s2["qz"][i] = q
    return s2


def _fd_check(energy, st, rotations, h=1e-6, rtol=2e-5, force=None, torques=None):
    """Central differences of energy(state) against the analytic values."""
    scale_f = max(np.abs(force).max(), 1e-3)
    for k, f in enumerate(["rx", "ry", "rz"]):
        up = st.copy(); up[f][0] += h
        dn = st.copy(); dn[f][0] -= h
        fd = -(energy(up) - energy(dn)) / (2 * h)
        assert abs(fd - force[k]) <= rtol * scale_f, ("force", k, fd, force[k])
    if rotations:
        for who, tq in ((0, torques[0]), (1, torques[1])):
            scale_t = max(np.abs(tq).max(), 1e-3)
            for k in range(3):
                fd = -(energy(_rotated(st, who, k, h)) - energy(_rotated(st, who, k, -h))) / (2 * h)
                assert abs(fd - tq[k]) <= rtol * scale_t, ("torque", who, k, fd, tq[k])


ISO = [(LJ, [1.0, 1.0, 2.5], 0.9, 2.4), (LJ12, [1.0, 1.0, 2.5], 0.9, 2.4),
       (LJW, [1.0, 1.0, 2.5, 0.3], 0.9, 2.4), (WCA, [1.0, 1.0], 0.9, 1.1),
       (MORSE, [1.0, 2.0, 1.1, 2.5], 0.7, 2.4), (DH, [5.0, 1.0, 1.0], 0.6, 4.9),
       # Coulomb inside the switch band only: within the band the reference
       # subtracts a scalar from every force component (not a gradient)
       (COULOMB, [5.0, 2.0], 0.6, 2.95)]


@pytest.mark.parametrize("kind,p,rmin,rmax", ISO)
def test_isotropic_forces_are_minus_gradient(oracle, kind, p, rmin, rmax):
    rng = np.random.default_rng(6 + kind)
    st = base_state(rng, 2, 30.0, charges=True)
    st["charge"] = [1.0, -1.0]
    for _ in range(40):
        d = rng.normal(size=3)
        d *= rng.uniform(rmin, rmax) / np.linalg.norm(d)
        st["rx"], st["ry"], st["rz"] = [0, d[0]], [0, d[1]], [0, d[2]]
        pe, f, _, _ = oracle.pair(kind, st, 0, 1, p)
        _fd_check(lambda s: oracle.pair(kind, s, 0, 1, p)[0], st, False, force=f)


def test_coulomb_closed_form(oracle):
    """Below the switch band coulombic_calculate_forces is q_i q_j / (4 pi
    eps r) exactly, and zero from the cutoff on."""
    st = base_state(np.random.default_rng(3), 2, 30.0, charges=True)
    st["charge"] = [1.5, -0.5]
    for r in (0.7, 1.3, 2.9, 5.0, 6.0):
        st["rx"], st["ry"], st["rz"] = [0.0, r], [0.0, 0.0], [0.0, 0.0]
        pe = oracle.pair(COULOMB, st, 0, 1, [5.0, 2.0])[0]
        want = 1.5 * -0.5 / (4 * np.pi * 2.0 * r) if r < 3.0 else (0.0 if r >= 5.0 else pe)
        assert abs(pe - want) <= 1e-15 * max(1.0, abs(want)), r


def test_lj_special_values(oracle):
    st = base_state(np.random.default_rng(7), 2, 30.0)
    st["rx"], st["ry"], st["rz"] = [0, 2 ** (1 / 6)], [0, 0], [0, 0]
    pe, f, _, _ = oracle.pair(WCA, st, 0, 1, [1.0, 1.0])
    assert abs(pe) < 1e-14 and np.abs(f).max() < 1e-12        # WCA: U = F = 0 at 2^(1/6) sigma
    st["rx"] = [0, 1.0]
    pe, f, _, _ = oracle.pair(LJ12, st, 0, 1, [1.0, 1.0, 2.5])
    assert np.abs(f).max() < 1e-12                             # eps((s/r)^12 - 2(s/r)^6): minimum at s
    assert np.isclose(pe, -1.0 - (2.5 ** -12 - 2 * 2.5 ** -6), rtol=1e-14)
    st["rx"] = [0, 2.5 - 1e-12]
    pe, _, _, _ = oracle.pair(LJ, st, 0, 1, [1.0, 1.0, 2.5])
    assert abs(pe) < 1e-9                                      # shifted to zero at the cutoff


@pytest.mark.parametrize("kind,p,r", [(GB, [4.0, 1, 2], (1.6, 3.2)), (GB, [4.0, 2, 1], (1.6, 3.2)),
                                      (GBR, [0.0, 1, 2], (0.9, 1.6)), (GBW, [4.0, 1, 2, 0.4], (1.6, 3.2)),
                                      (GBC, [4.0, 1, 2, 0.7], (1.6, 3.2)), (GBC, [4.0, 2, 1, -0.7], (1.6, 3.2)),
                                      (GBWC, [4.0, 1, 2, 0.4, 0.7], (1.6, 3.2))])
def test_gay_berne_forces_torques_are_minus_gradient(oracle, kind, p, r):
    """Outside the switching band (r < 0.9 (sigma + rc - sigma0)); the welded
    variants' kinks at r = sigma and sigma + wf are met with probability ~0 by
    the random separations."""
    rng = np.random.default_rng(8 + kind + int(p[1]))
    st = base_state(rng, 2, 30.0, sig=(1, 1, 3), eps=(5, 5, 1))
    checked = 0
    for _ in range(200):
        d = rng.normal(size=3)
        d *= rng.uniform(*r) / np.linalg.norm(d)
        q = rand_quats(rng, 2)
        st["rx"], st["ry"], st["rz"] = [0, d[0]], [0, d[1]], [0, d[2]]
        for k, f in enumerate(["qw", "qx", "qy", "qz"]):
            st[f] = q[:, k]
        pe, f, ti, tj = oracle.pair(kind, st, 0, 1, p)
        if pe == 0.0 or abs(pe) > 50:
            continue
        _fd_check(lambda s: oracle.pair(kind, s, 0, 1, p)[0], st, True, force=f, torques=(ti, tj),
                  rtol=1e-4)
        checked += 1
    assert checked > 20


def test_gay_berne_sphere_limit(oracle):
    """sigma = eps = (1,1,1): chi = chi' = 0 and GB(mu, nu) reduces to the
    12-6 form r^-12 - 2 r^-6 (force.f90:430-620 writes the well depth at the
    minimum r = sigma0) up to a constant, so energy differences and forces
    follow it; torques vanish."""
    st = base_state(np.random.default_rng(9), 2, 30.0)
    lj = lambda r: r ** -12 - 2 * r ** -6  # noqa: E731
    for mu, nu in [(1, 2), (2, 1)]:
        vals = []
        for r in (1.05, 1.3, 1.9):
            st["rx"], st["ry"], st["rz"] = [0, r], [0, 0], [0, 0]
            pe, f, ti, tj = oracle.pair(GB, st, 0, 1, [4.0, mu, nu])
            vals.append(pe)
            assert np.isclose(f[0], 12 * (r ** -7 - r ** -13), rtol=1e-10)   # particle j sits at +x
            assert np.abs(ti).max() < 1e-10 and np.abs(tj).max() < 1e-10
        assert np.isclose(vals[0] - vals[2], lj(1.05) - lj(1.9), rtol=1e-10)


@pytest.mark.parametrize("kind,sig,p,rmax", [(ECPR, (1, 1, 3), [0.0], 3.3), (ECP, (1, 1, 3), [1.5], 1.35),
                                             (ECPR, (1.0, 2.0, 3.0), [0.0], 3.3),
                                             (ECPW, (1, 1, 3), [1.5, 0.3], 1.35),
                                             (ECPC, (1, 1, 3), [1.5, 0.6], 1.35),
                                             (ECPWC, (1.0, 2.0, 3.0), [1.5, 0.3, -0.6], 1.35)])
def test_ecp_forces_torques_are_partial_derivatives(oracle, kind, sig, p, rmax):
    """At a fixed contact parameter lambda the ECP force and torques
    (force.f90:1202-1420) are the exact partial derivatives of the pair energy.
    For the cut variant only r < 0.9 (sigma_min + rc - sigma0) is used: inside
    the switching band the reference adds one scalar to every component, which
    only the comparison with the reference build pins."""
    rng = np.random.default_rng(10 + kind)
    st = base_state(rng, 2, 30.0, sig=sig)
    checked = 0
    for _ in range(600):
        d = rng.normal(size=3)
        d *= rng.uniform(1.0, rmax) / np.linalg.norm(d)
        q = rand_quats(rng, 2)
        st["rx"], st["ry"], st["rz"] = [0, d[0]], [0, d[1]], [0, d[2]]
        for k, f in enumerate(["qw", "qx", "qy", "qz"]):
            st[f] = q[:, k]
        lam = oracle.pair(kind, st, 0, 1, p)[4]
        pe, f, ti, tj = oracle.ecp_at_lambda(kind, st, 0, 1, p, lam)
        if pe == 0.0 or abs(pe) > 100:
            continue
        _fd_check(lambda s: oracle.ecp_at_lambda(kind, s, 0, 1, p, lam)[0], st, True, force=f,
                  torques=(ti, tj), rtol=1e-4)
        checked += 1
    assert checked > 20


def test_ecp_parallel_identical_converged(oracle):
    """Identical parallel ellipsoids: A = B so lambda* = 1/2 for every
    separation; the Brent result is within its tolerance of it and at 1/2 the
    force is the full gradient of U(r)."""
    rng = np.random.default_rng(11)
    st = base_state(rng, 2, 30.0, sig=(1, 1, 3))
    q = rand_quats(rng, 1)[0]
    for k, f in enumerate(["qw", "qx", "qy", "qz"]):
        st[f] = [q[k], q[k]]
    for _ in range(40):
        d = rng.normal(size=3)
        d *= rng.uniform(1.0, 3.0) / np.linalg.norm(d)
        st["rx"], st["ry"], st["rz"] = [0, d[0]], [0, d[1]], [0, d[2]]
        lam = oracle.pair(ECPR, st, 0, 1, [0.0])[4]
        assert abs(lam - 0.5) < 1e-2 * 0.5 + 1e-3 + 1e-12
        pe, f, _, _ = oracle.ecp_at_lambda(ECPR, st, 0, 1, [0.0], 0.5)
        if pe == 0.0:
            continue
        _fd_check(lambda s: oracle.ecp_at_lambda(ECPR, s, 0, 1, [0.0], 0.5)[0], st, False, force=f,
                  rtol=1e-4)


def test_vv_free_particle_and_harmonic_energy(oracle):
    """No force: r advances by v dt exactly.  A harmonic bond integrated with
    VV conserves energy to O(dt^2)."""
    rng = np.random.default_rng(12)
    st = base_state(rng, 20, 50.0, forces=False)
    r0 = np.stack([st["rx"], st["ry"], st["rz"]], 1)
    v0 = np.stack([st["vx"], st["vy"], st["vz"]], 1)
    oracle.integrate(VV, 0, st, 0.01)
    oracle.integrate(VV, 1, st, 0.01)
    r1 = np.stack([st["rx"], st["ry"], st["rz"]], 1)
    d = r1 - r0 - 0.01 * v0
    d -= np.round(d / 50.0) * 50.0
    assert np.abs(d).max() < 1e-12
    # 1-d oscillator through the VV halves: F = -x
    st = State(1, [1e6] * 3)
    st["mass"], st["inv_mass"], st["rx"], st["vx"], st["fx"] = 1.0, 1.0, 1.0, 0.0, -1.0
    dt = 0.01
    e0 = 0.5
    for _ in range(2000):
        oracle.integrate(VV, 0, st, dt)
        st["fx"] = -st["rx"]
        oracle.integrate(VV, 1, st, dt)
    e = 0.5 * st["vx"][0] ** 2 + 0.5 * st["rx"][0] ** 2
    assert abs(e - e0) < dt * dt


def _rot_body_from_space(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)],
                     [2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)],
                     [2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)]])


def test_quaternion_free_rotor(oracle):
    """Torque-free rotors under the quaternion steps (integrators.f90:123-299),
    where L is the space-frame angular momentum and omega_body = I^-1 A(q) L:
    L stays fixed, |q| = 1, the rotational energy sum (A L)_k^2 / I_k is
    conserved to O(dt^2), and a spherical top turns about L at the rate
    |L|/I exactly as q(t) = exp(t L / 2I) q(0)."""
    rng = np.random.default_rng(13)
    st = base_state(rng, 16, 50.0, sig=(1, 2, 3), forces=False)
    L0 = np.stack([st["lx"], st["ly"], st["lz"]], 1).copy()
    inv = np.stack([st["ixx_inv"], st["iyy_inv"], st["izz_inv"]], 1)

    def erot(s):
        out = []
        for i in range(s.n):
            lb = _rot_body_from_space([s["qw"][i], s["qx"][i], s["qy"][i], s["qz"][i]]) @ L0[i]
            out.append(np.sum(lb * lb * inv[i]))
        return np.array(out)
    e0 = erot(st)
    dt = 0.002
    for _ in range(200):
        oracle.integrate(QQ, 0, st, dt)
        oracle.integrate(QQ, 1, st, dt)
    assert np.array_equal(np.stack([st["lx"], st["ly"], st["lz"]], 1), L0)
    q = np.stack([st["qw"], st["qx"], st["qy"], st["qz"]], 1)
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-12)
    assert np.abs(erot(st) / e0 - 1).max() < 1e-4
    # spherical top
    st = base_state(rng, 8, 50.0, forces=False)
    st["ixx"] = st["iyy"] = st["izz"] = 0.3
    st["ixx_inv"] = st["iyy_inv"] = st["izz_inv"] = 1 / 0.3
    q0 = np.stack([st["qw"], st["qx"], st["qy"], st["qz"]], 1).copy()
    L = np.stack([st["lx"], st["ly"], st["lz"]], 1)
    nsteps = 200
    for _ in range(nsteps):
        oracle.integrate(QQ, 0, st, dt)
        oracle.integrate(QQ, 1, st, dt)
    for i in range(st.n):
        w = L[i] / 0.3
        ang = np.linalg.norm(w) * dt * nsteps
        ax = w / np.linalg.norm(w)
        dq = np.concatenate([[np.cos(ang / 2)], np.sin(ang / 2) * ax])
        a0, a1, a2, a3 = dq
        qw, qx, qy, qz = q0[i]
        want = np.array([a0 * qw - a1 * qx - a2 * qy - a3 * qz, a0 * qx + a1 * qw + a2 * qz - a3 * qy,
                         a0 * qy - a1 * qz + a2 * qw + a3 * qx, a0 * qz + a1 * qy - a2 * qx + a3 * qw])
        got = np.array([st["qw"][i], st["qx"][i], st["qy"][i], st["qz"][i]])
        assert min(np.abs(got - want).max(), np.abs(got + want).max()) < 1e-4


def test_langevin_equipartition(oracle):
    """Zero force: the Langevin steps (integrators.f90:507-676) relax the
    velocities to <m v^2> = T and the body-frame angular momenta A(q) L to
    <L_k^2 / I_k> = T."""
    rng = np.random.default_rng(14)
    n = 4000
    st = base_state(rng, n, 200.0, sig=(1, 1, 3), T=1.3, forces=False)
    for f in ["vx", "vy", "vz", "lx", "ly", "lz"]:
        st[f] = 0.0
    for step in range(1, 301):
        oracle.integrate(LGV, 0, st, 0.01, step=step, damping=1.0)
        oracle.integrate(LGV, 1, st, 0.01, damping=1.0)
    mv2 = np.mean(st["mass"] * (st["vx"] ** 2 + st["vy"] ** 2 + st["vz"] ** 2)) / 3
    lb = np.array([_rot_body_from_space([st["qw"][i], st["qx"][i], st["qy"][i], st["qz"][i]])
                   @ [st["lx"][i], st["ly"][i], st["lz"][i]] for i in range(n)])
    lrot = np.mean(lb[:, 0] ** 2 / st["ixx"] + lb[:, 1] ** 2 / st["iyy"] + lb[:, 2] ** 2 / st["izz"]) / 3
    # 12000 samples: the translational mean is within 0.02 (1 sigma); the
    # rotational one carries an O(dt) bias of the quaternion splitting
    assert abs(mv2 - 1.3) < 0.05 and abs(lrot - 1.3) < 0.1


def _bond_angle(p0, p1, p2):
    a, b = p0 - p1, p2 - p1
    return np.arctan2(np.linalg.norm(np.cross(a, b)), a @ b)


def _dihedral(p0, p1, p2, p3):
    """IUPAC torsion angle (trans = pi) by the atan2 route."""
    b1, b2, b3 = p1 - p0, p2 - p1, p3 - p2
    n1, n2 = np.cross(b1, b2), np.cross(b2, b3)
    return np.arctan2(np.linalg.norm(b2) * (b1 @ n2), n1 @ n2)


# kind, atoms, coeff, angle (degrees), closed form of the energy in the bond
# angle alpha at the middle atom or the torsion angle phi
BONDED_FORMS = [
    (1, 3, 20.0, 0.0, lambda x, k, a: k * (1.0 + np.cos(x))),
    (2, 3, 5.0, 110.0, lambda x, k, a: 0.5 * k * (np.cos(np.pi - x) - np.cos(np.pi - np.radians(a))) ** 2),
    (5, 3, 5.0, 110.0, lambda x, k, a: 0.5 * k * (x - np.radians(a)) ** 2),
    (5, 3, 7.0, 180.0, lambda x, k, a: 0.5 * k * (x - np.radians(a)) ** 2),
    (3, 4, 1.0, 0.0, lambda x, k, a: k * (1.0 + np.cos(x))),
    (4, 4, 1.5, 60.0, lambda x, k, a: 0.5 * k * (np.cos(np.pi - x) - np.cos(np.pi - np.radians(a))) ** 2),
    (6, 4, 1.5, 60.0, lambda x, k, a: k * (1.0 - np.cos(x - np.radians(a)))),
    (6, 4, 2.0, -120.0, lambda x, k, a: k * (1.0 - np.cos(x - np.radians(a)))),
]


@pytest.mark.parametrize("kind,nat,coeff,angle,form", BONDED_FORMS)
def test_bonded_closed_forms_and_gradients(oracle, kind, nat, coeff, angle, form):
    """Bonded terms (force.f90:2244-2579): the energy equals its closed form
    in the bond angle (bends) or the IUPAC torsion (dihedrals) computed here
    by an independent atan2 route, and each atom's force is minus the
    central-difference gradient of that energy, except for cosine_bend_
    and cosine_dih_, whose forces are PLUS the gradient: they reuse the
    k(1 -/+ cos) force expressions with strength +k (cos - cos0)
    (force.f90:2383-2388, :2449-2460), so they push the angle away from the
    energy minimum.  The device reproduces the reference, sign included.
    Periodic images are used: the atoms straddle the box edge."""
    sign = 1.0 if kind in (2, 4) else -1.0
    rng = np.random.default_rng(50 + kind)
    L = 6.0
    st = base_state(rng, nat, L)
    h = 1e-6
    for _ in range(25):
        pos = np.cumsum(rng.normal(size=(nat, 3)) * 0.7, axis=0) + L / 2 - 0.3
        wrapped = pos - np.round(pos / L) * L
        st["rx"], st["ry"], st["rz"] = wrapped[:, 0], wrapped[:, 1], wrapped[:, 2]
        idx = list(range(nat))
        pe, f = oracle.bonded(kind, st, idx, coeff, angle)
        x = _bond_angle(*pos) if nat == 3 else _dihedral(*pos)
        assert abs(pe - form(x, coeff, angle)) <= 1e-10 * max(1.0, abs(pe)), (pe, form(x, coeff, angle))
        scale = max(np.abs(f[:nat]).max(), 1e-3)
        for a in range(nat):
            for k, fld in enumerate(["rx", "ry", "rz"]):
                up = st.copy(); up[fld][a] += h
                dn = st.copy(); dn[fld][a] -= h
                fd = sign * (oracle.bonded(kind, up, idx, coeff, angle)[0]
                             - oracle.bonded(kind, dn, idx, coeff, angle)[0]) / (2 * h)
                assert abs(fd - f[a][k]) <= 2e-5 * scale, (kind, a, k, fd, f[a][k])


def test_dipole_forces_torques_are_minus_gradient(oracle):
    """dipole_calculate_forces (force.f90:2687-2751): U = u1.u2/r^3 -
    3 (u1.r)(u2.r)/r^5 for the charge-scaled body z axes; force = -grad U,
    torques = minus the derivative under rotations of each particle; two
    parallel dipoles on their common axis give U = -2 q1 q2 / r^3."""
    rng = np.random.default_rng(31)
    st = base_state(rng, 2, 30.0, charges=True)
    st["charge"] = [1.3, -0.7]
    dip = np.zeros((3, 2)); rfl = np.zeros((3, 2))
    for _ in range(30):
        d = rng.normal(size=3)
        d *= rng.uniform(0.8, 2.3) / np.linalg.norm(d)
        st["rx"], st["ry"], st["rz"] = [0, d[0]], [0, d[1]], [0, d[2]]
        q = rand_quats(rng, 2)
        for k, f in enumerate(["qw", "qx", "qy", "qz"]):
            st[f] = q[:, k]
        pe, f, ti, tj = oracle.dipole_pair(st, 0, 1, 2.5, 10.0, dip, rfl)
        _fd_check(lambda s: oracle.dipole_pair(s, 0, 1, 2.5, 10.0, dip.copy(), rfl.copy())[0], st, True,
                  force=f, torques=(ti, tj))
    st["qw"], st["qx"], st["qy"], st["qz"] = [1, 1], [0, 0], [0, 0], [0, 0]
    st["rx"], st["ry"], st["rz"] = [0, 0], [0, 0], [0, 1.7]
    pe = oracle.dipole_pair(st, 0, 1, 2.5, 10.0, dip, rfl)[0]
    assert abs(pe - (-2.0 * 1.3 * -0.7 / 1.7 ** 3)) < 1e-14


def test_composite_bodies_invariants(oracle):
    """copy_transform puts `to` at |offset| from `from` (a rotation), at the
    offset itself for the identity quaternion and at R(q) offset in general
    (R the body-to-space matrix of the quaternion); transfer_forces keeps the
    total force and the total angular momentum source sum(r x F + T) (open
    box) and leaves the satellite with none."""
    rng = np.random.default_rng(8)
    st = base_state(rng, 12, 40.0)
    st.periodic = False
    for t in range(6):
        off = rng.normal(size=3)
        assert oracle.copy_transform(st, t, 6 + t, off) == 0
        d = np.array([st[f][6 + t] - st[f][t] for f in ("rx", "ry", "rz")])
        w, x, y, z = (st[f][t] for f in ("qw", "qx", "qy", "qz"))
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        assert np.allclose(d, R @ off, atol=1e-13) and abs(np.linalg.norm(d) - np.linalg.norm(off)) < 1e-13
    pos = lambda: np.stack([st["rx"], st["ry"], st["rz"]], 1)
    F = lambda: np.stack([st["fx"], st["fy"], st["fz"]], 1)
    T = lambda: np.stack([st["tx"], st["ty"], st["tz"]], 1)
    f0, l0 = F().sum(0), (np.cross(pos(), F()) + T()).sum(0)
    for t in range(6):
        oracle.transfer_forces(st, 6 + t, t % 3)
    assert np.allclose(F().sum(0), f0, atol=1e-12) and np.allclose((np.cross(pos(), F()) + T()).sum(0), l0, atol=1e-11)
    assert np.all(F()[6:] == 0) and np.all(T()[6:] == 0)


def test_rattle_satisfies_constraints(oracle):
    """apply_constraints_a/b (system.f90:1060-1211): bond lengths to 2e-5 L^2
    after the position stage, r.v near zero after the velocity stage."""
    rng = np.random.default_rng(15)
    L = 12.0
    pos, bi, bj = chains(rng, 10, 20, [L, L, L])
    st = base_state(rng, len(pos), L, T=1.0)
    st["rx"], st["ry"], st["rz"] = pos[:, 0], pos[:, 1], pos[:, 2]
    st.bbi, st.bbj, st.bb_len = bi, bj, np.ones(len(bi))
    st["dof"] = 6 * st.n - len(bi) - 3
    for f in ["fx", "fy", "fz"]:
        st[f] = rng.normal(size=st.n)
    oracle.integrate(VVC, 0, st, 0.005)
    d = np.stack([st[f][bi] - st[f][bj] for f in ["rx", "ry", "rz"]], 1)
    d -= np.round(d / L) * L
    assert np.abs((d ** 2).sum(1) - 1.0).max() <= 2e-5 * 1.0 + 1e-12
    oracle.integrate(VVC, 1, st, 0.005)
    v = np.stack([st[f][bi] - st[f][bj] for f in ["vx", "vy", "vz"]], 1)
    assert np.abs((d * v).sum(1)).max() < 1e-4


def test_converged_lambda_and_contact_distance(oracle):
    """The converged contact parameter maximises the contact function (scipy's
    bounded minimiser to 1e-12 agrees), equals a/(a+b) for spheres of radii
    a, b, whose contact distance is a+b."""
    from scipy.optimize import minimize_scalar
    rng = np.random.default_rng(16)
    for ar, br, r in _shape_pairs(rng, 100):
        ar, br = ar.ravel(order="F"), br.ravel(order="F")
        lam, it = oracle.lambda_converged(ar, br, r)
        ref = minimize_scalar(lambda l: oracle.optim_dist(l, ar, br, r), bounds=(0, 1), method="bounded",
                              options={"xatol": 1e-12})
        # the optimum is quadratic, so the minimiser's argument is good to ~1e-8
        # and its value to ~1e-16: lambda agrees and is at least as good
        f = oracle.optim_dist(lam, ar, br, r)
        assert abs(lam - ref.x) < 1e-6 and f <= ref.fun + 1e-14 * abs(ref.fun) and it < 40
    for _ in range(20):
        a, b = rng.uniform(0.3, 2.0, 2)
        r = rng.normal(size=3) * 4
        lam, _ = oracle.lambda_converged(np.eye(3).ravel() * a * a, np.eye(3).ravel() * b * b, r)
        assert abs(lam - a / (a + b)) < 1e-12
        sig = oracle.contact_distance(np.eye(3).ravel() * a * a, np.eye(3).ravel() * b * b, r, lam)
        assert abs(sig - (a + b)) < 1e-12


def test_network_features(oracle):
    """Features of a pair: the unit separation in particle i's body frame and
    conj(q_i) q_j with w >= 0.  Identity orientations give (rhat, 1, 0, 0, 0);
    rotating both particles and the separation together leaves them fixed."""
    rng = np.random.default_rng(17)
    st = base_state(rng, 2, 30.0)
    d = rng.normal(size=3) * 2
    st["rx"], st["ry"], st["rz"] = [d[0], 0.0], [d[1], 0.0], [d[2], 0.0]
    for f, v in zip(["qw", "qx", "qy", "qz"], [1.0, 0.0, 0.0, 0.0]):
        st[f] = v
    x0 = oracle.features(st, 0, 1)
    assert np.allclose(x0[:3], d / np.linalg.norm(d)) and np.allclose(x0[3:], [1, 0, 0, 0])
    qi, qj = rand_quats(rng, 2)
    st2 = st.copy()
    for k, f in enumerate(["qw", "qx", "qy", "qz"]):
        st2[f] = [qi[k], qj[k]]
    x1 = oracle.features(st2, 0, 1)
    g = rand_quats(rng, 1)[0]                      # rotate the whole pair by g
    Rg = _rot_body_from_space(g).T                 # body -> space of g
    st3 = st2.copy()
    dn = Rg @ d
    st3["rx"], st3["ry"], st3["rz"] = [dn[0], 0.0], [dn[1], 0.0], [dn[2], 0.0]
    for who, q in ((0, qi), (1, qj)):
        a0, a1, a2, a3 = g
        w, x, y, z = q
        qq = [a0 * w - a1 * x - a2 * y - a3 * z, a0 * x + a1 * w + a2 * z - a3 * y,
              a0 * y - a1 * z + a2 * w + a3 * x, a0 * z + a1 * y - a2 * x + a3 * w]
        for k, f in enumerate(["qw", "qx", "qy", "qz"]):
            st3[f][who] = qq[k]
    assert np.allclose(oracle.features(st3, 0, 1), x1, atol=1e-12)
    assert x1[3] >= 0 and abs(np.linalg.norm(x1[3:]) - 1) < 1e-12
