"""The C oracle vs the SUCROSE reference compiled from its own Fortran sources.

Both sides run on identical FP64 inputs; every comparison is bit-for-bit
(``np.array_equal``), because the oracle restates the reference's expressions
in the reference's evaluation order without contraction.  These tests skip
when ``oracle/_ref`` has not been built (``make -C oracle ref``).
"""
import ctypes as C

import numpy as np
import pytest

from oracle.oracle import (COULOMB, DH, ECP, ECPC, ECPR, ECPW, ECPWC, GB, GBC, GBR, GBW, GBWC, LGV, LGVC, LJ, LJ12, LJW, MORSE,
                           NHT, NHTC, QQ, VV, VVC, WCA, State, reference_available)
from tests._states import base_state, chains

pytestmark = [pytest.mark.ref,
              pytest.mark.skipif(not reference_available(), reason="reference not built (make -C oracle ref)")]


def same_state(a: State, b: State):
    return np.array_equal(a.buf, b.buf) and np.array_equal(a.scal, b.scal)


def test_philox_uniform_words(oracle, reference):
    rng = np.random.default_rng(0)
    ctrs = rng.integers(-2**31, 2**31, size=(4096, 4))
    ctrs[:8] = [[0, 0, 0, 0], [-1, -1, -1, -1], [1, 2, 3, 4], [2**31 - 1] * 4,
                [-2**31] * 4, [7, 0, 0, 1000], [123456, 789, 1, 0], [0, -5, 99, 3]]
    import ctypes as C
    for c in ctrs:
        a = np.zeros(4)
        oracle.lib.or_philox_uniform((C.c_int32 * 4)(*[int(x) for x in c]), a.ctypes.data_as(C.POINTER(C.c_double)))
        assert np.array_equal(a, reference.philox_uniform(c)), c


def test_philox_normals(oracle, reference):
    rng = np.random.default_rng(1)
    for c in rng.integers(-2**31, 2**31, size=(2048, 4)):
        assert np.array_equal(oracle.philox_normal(*[int(x) for x in c]),
                              reference.philox_normal(*[int(x) for x in c])), c


def test_cholesky_solve(oracle, reference):
    import ctypes as C
    rng = np.random.default_rng(2)
    for _ in range(500):
        m = rng.normal(size=(3, 3))
        a = m @ m.T + 3 * np.eye(3)
        b = rng.normal(size=3)
        o = np.zeros(3)
        af = np.asfortranarray(a)
        oracle.lib.or_cholesky_solve(af.ctypes.data_as(C.POINTER(C.c_double)),
                                     b.ctypes.data_as(C.POINTER(C.c_double)),
                                     o.ctypes.data_as(C.POINTER(C.c_double)))
        assert np.array_equal(o, reference.cholesky_solve(a, b))


def _pair_state(rng, npairs, kw, rmin, rmax, box=40.0):
    """npairs independent pairs (2k, 2k+1) at distances in [rmin, rmax);
    a quarter of them straddle the periodic boundary."""
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


PAIRS = [
    (LJ, [1.0, 1.0, 2.5], dict(), 0.8, 2.8),
    (LJ12, [1.0, 1.0, 2.5], dict(), 0.8, 2.8),
    (LJW, [1.0, 1.0, 2.5, 0.3], dict(), 0.8, 2.8),
    (WCA, [1.0, 1.0], dict(), 0.8, 1.3),
    (MORSE, [1.0, 2.0, 1.1, 2.5], dict(), 0.6, 2.8),
    (DH, [5.0, 1.0, 1.0], dict(charges=True), 0.5, 5.5),
    (GB, [4.0, 1, 2], dict(sig=(1, 1, 3), eps=(5, 5, 1)), 0.7, 4.5),
    (GBR, [0.0, 1, 2], dict(sig=(1, 1, 3), eps=(5, 5, 1)), 0.7, 4.5),
    (ECP, [1.5], dict(sig=(1, 1, 3)), 0.6, 3.4),
    (ECPR, [0.0], dict(sig=(1, 1, 3)), 0.6, 3.4),
    (ECPR, [0.0], dict(sig=(1.0, 2.0, 3.0)), 0.6, 3.4),
    # welded and chiral variants (force.f90:535-1030, 1553-2234); GB (mu, nu)
    # fixed at the reference's (1, 2)
    (GBW, [4.0, 1, 2, 0.4], dict(sig=(1, 1, 3), eps=(5, 5, 1)), 0.7, 4.9),
    (GBC, [4.0, 1, 2, 0.7], dict(sig=(1, 1, 3), eps=(5, 5, 1)), 0.7, 4.5),
    (GBWC, [4.0, 1, 2, 0.4, -0.7], dict(sig=(1, 1, 3), eps=(5, 5, 1)), 0.7, 4.9),
    (ECPW, [1.5, 0.3], dict(sig=(1, 1, 3)), 0.6, 3.7),
    (ECPC, [1.5, 0.6], dict(sig=(1, 1, 3)), 0.6, 3.4),
    (ECPWC, [1.5, 0.3, -0.6], dict(sig=(1.0, 2.0, 3.0)), 0.6, 3.7),
    (COULOMB, [5.0, 2.0], dict(charges=True), 0.5, 5.5),
]


@pytest.mark.parametrize("kind,p,kw,rmin,rmax", PAIRS)
def test_pair_kernels_bit_exact(oracle, reference, kind, p, kw, rmin, rmax):
    rng = np.random.default_rng(10 + kind)
    npairs = 600
    st = _pair_state(rng, npairs, kw, rmin, rmax)
    reference.load(st)
    nonzero = 0
    for k in range(npairs):
        i, j = 2 * k, 2 * k + 1
        a = oracle.pair(kind, st, i, j, p)[:4]
        b = reference.pair(kind, i, j, p)
        nonzero += b[0] != 0.0
        assert a[0] == b[0] and all(np.array_equal(x, y) for x, y in zip(a[1:], b[1:])), (kind, k)
    assert nonzero > npairs // 10


def test_ecp_brent_bit_exact(oracle, reference):
    import ctypes as C
    from tests._states import rand_quats
    rng = np.random.default_rng(4)
    PD = C.POINTER(C.c_double)
    for _ in range(400):
        q = rand_quats(rng, 2)
        sem = [np.diag(rng.uniform(0.3, 2.0, 3) ** -2) for _ in range(2)]
        mats = []
        for qq, s in zip(q, sem):
            w, x, y, z = qq
            A = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)],
                          [2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)],
                          [2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)]])
            mats.append(np.asfortranarray(A.T @ s @ A))
        r = rng.normal(size=3) * rng.uniform(0.5, 4.0)
        nit = C.c_int()
        lam = oracle.lib.or_ecp_brent(mats[0].ctypes.data_as(PD), mats[1].ctypes.data_as(PD),
                                      r.ctypes.data_as(PD), C.byref(nit))
        assert lam == reference.brent(mats[0], mats[1], r)


BONDED = [(1, 3, 20.0, 0.0), (2, 3, 20.0, np.pi), (2, 3, 5.0, 2.0), (3, 4, 1.0, 0.0), (4, 4, 1.0, 0.7),
          (5, 3, 5.0, 110.0), (5, 3, 20.0, 0.0), (6, 4, 1.0, 60.0), (6, 4, 2.0, 0.0)]


@pytest.mark.parametrize("kind,nat,coeff,angle", BONDED)
def test_bonded_bit_exact(oracle, reference, kind, nat, coeff, angle):
    rng = np.random.default_rng(20 + kind)
    pos, _, _ = chains(rng, 60, nat, [9.0, 9.0, 9.0], bond=1.0)
    st = base_state(rng, len(pos), 9.0)
    st["rx"], st["ry"], st["rz"] = pos[:, 0], pos[:, 1], pos[:, 2]
    reference.load(st)
    for c in range(60):
        idx = list(range(c * nat, c * nat + nat))
        a = oracle.bonded(kind, st, idx, coeff, angle)
        b = reference.bonded(kind, idx, coeff, angle)
        assert a[0] == b[0] and np.array_equal(a[1][:nat], b[1][:nat]), (kind, c)


INTEGRATORS = [(VV, None), (QQ, None), (NHT, None), (NHT, 2.5), (LGV, None)]


@pytest.mark.parametrize("kind,coupling", INTEGRATORS)
@pytest.mark.parametrize("phase", [0, 1])
def test_integrator_one_phase(oracle, reference, kind, coupling, phase):
    rng = np.random.default_rng(30 + kind + 7 * phase)
    st = base_state(rng, 96, 6.0, sig=(1, 1, 3), eps=(5, 5, 1), T=1.5)
    st["peta0"], st["peta1"], st["peta2"] = 0.3, -0.2, 0.1
    st["eta0"], st["eta1"], st["eta2"] = 0.05, 0.01, -0.02
    st["kin_eng"], st["kin_rot"] = 70.0, 60.0
    reference.load(st)
    reference.integrate(kind, phase, 0.003, step=7, damping=1.3, coupling=coupling)
    b = st.copy()
    reference.store(b)
    a = st.copy()
    oracle.integrate(kind, phase, a, 0.003, step=7, damping=1.3, coupling=coupling)
    assert same_state(a, b)


def _chain_state(rng, nch=6, nb=12, L=9.0):
    pos, bi, bj = chains(rng, nch, nb, [L, L, L])
    st = base_state(rng, len(pos), L, T=1.0)
    st["rx"], st["ry"], st["rz"] = pos[:, 0], pos[:, 1], pos[:, 2]
    st.bbi, st.bbj, st.bb_len = bi, bj, np.ones(len(bi))
    st["dof"] = 6 * st.n - len(bi) - 3
    return st


@pytest.mark.parametrize("kind", [VVC, LGVC, NHTC])
@pytest.mark.parametrize("phase", [0, 1])
def test_constrained_integrators(oracle, reference, kind, phase):
    st = _chain_state(np.random.default_rng(40 + kind + phase))
    reference.load(st)
    reference.integrate(kind, phase, 0.005, step=3, damping=1.0)
    b = st.copy()
    reference.store(b)
    a = st.copy()
    oracle.integrate(kind, phase, a, 0.005, step=3, damping=1.0)
    assert same_state(a, b)


@pytest.mark.parametrize("phase", [0, 1])
def test_rattle_bit_exact(oracle, reference, phase):
    import ctypes as C
    st = _chain_state(np.random.default_rng(50 + phase))
    # perturb the bonds so that the solver iterates
    for f in ["rx", "ry", "rz", "vx", "vy", "vz"]:
        st[f] = st[f] + np.random.default_rng(5).normal(size=st.n) * 0.02
    st["rx_old"], st["ry_old"], st["rz_old"] = st["rx"] - 0.01, st["ry"] + 0.005, st["rz"]
    reference.load(st)
    reference.constraints(phase, 0.005)
    b = st.copy()
    reference.store(b)
    a = st.copy()
    oracle.call("or_apply_constraints_a" if phase == 0 else "or_apply_constraints_b", a, C.c_double(0.005))
    assert same_state(a, b)


def test_system_routines(oracle, reference):
    import ctypes as C
    rng = np.random.default_rng(60)
    st = base_state(rng, 128, 7.0)
    st["kin_eng"], st["kin_rot"], st["virial_bonded"] = 3.3, 1.1, 0.7
    for name, call in [("or_calculate_thermo", reference.calculate_thermo),
                       ("or_initialize_forces", reference.initialize_forces)]:
        reference.load(st)
        call()
        b = st.copy()
        reference.store(b)
        a = st.copy()
        oracle.call(name, a)
        assert same_state(a, b), name
    for rc in (0, 1):
        reference.load(st)
        reference.initialize_system(remove_com=bool(rc))
        b = st.copy()
        reference.store(b)
        a = st.copy()
        oracle.call("or_initialize_defaults", a, C.c_int(rc))
        assert same_state(a, b), ("initialize_system", rc)
        reference.load(st)
        reference.set_velocity_to_temperature(1.7, 15 + rc, remove_com=bool(rc))
        b = st.copy()
        reference.store(b)
        a = st.copy()
        oracle.call("or_set_velocity_to_temperature", a, C.c_double(1.7), C.c_int(15 + rc), C.c_int(rc))
        assert same_state(a, b), ("set_velocity_to_temperature", rc)


def test_fcc_lattice(oracle, reference):
    import ctypes as C
    st = State(4 * 6 ** 3, [1.0, 1.0, 1.0])
    reference.load(st)
    reference.generate_fcc_lattice(6, 0.8442)
    b = st.copy()
    reference.store(b)
    a = st.copy()
    oracle.call("or_generate_fcc_lattice", a, C.c_int(6), C.c_double(0.8442))
    assert np.array_equal(a.buf[:, :3], b.buf[:, :3])
    assert np.array_equal(a.buf[:, 9:13], b.buf[:, 9:13])   # quaternions
    box = np.zeros(3)
    reference.lib.ref_get_box(box.ctypes.data_as(C.POINTER(C.c_double)))
    assert np.array_equal(a.box, box) and a["density"] == b["density"]


NLIST_CASES = [
    (1, 300, [9.0, 9.0, 9.0], 2.5, 0.3),
    (2, 500, [12.0, 10.0, 14.0], 2.5, 0.3),
    (3, 200, [5.4, 5.4, 5.4], 2.5, 0.3),      # list sphere close to L/2
    (4, 864, [10.0776] * 3, 2.5, 0.3),
]


@pytest.mark.parametrize("seed,n,box,rc,skin", NLIST_CASES)
def test_nlist_and_rebuild(seed, n, box, rc, skin):
    from tests._refjobs import in_child, nlist_job
    res = in_child(nlist_job, seed, n, box, rc, skin, [0.01, 0.05, 0.2, 0.5])
    name, same, nnz = res[0]
    assert same and nnz > n
    for name, decision_equal, same, nnz, rebuilt in res[1:]:
        assert decision_equal and same, name
    rebuilt = [r[-1] for r in res[1:]]
    assert not rebuilt[0] and any(rebuilt)


RUNS = [
    (VV, LJ, [1.0, 1.0, 2.5], 256, 8.0, (1, 1, 1), (1, 1, 1), 2.5, 0.3, 30, 0.005, 1.0),
    (LGV, LJ, [1.0, 1.0, 2.5], 256, 8.0, (1, 1, 1), (1, 1, 1), 2.5, 0.3, 30, 0.005, 1.0),
    (NHT, GB, [4.0, 1, 2], 64, 14.0, (1, 1, 3), (5, 5, 1), 4.0, 0.4, 10, 0.0015, 3.0),
    (LGV, ECPR, [0.0], 64, 14.0, (1, 1, 3), (1, 1, 1), 3.1, 0.3, 10, 0.002, 1.0),
]


@pytest.mark.parametrize("integ,kind,p,n,box,sig,eps,rc,skin,nsteps,dt,T", RUNS)
def test_driver_loop_bit_exact(integ, kind, p, n, box, sig, eps, rc, skin, nsteps, dt, T):
    from tests._refjobs import in_child, run_job
    same, diff, pe, temp = in_child(run_job, integ, kind, p, n, box, 5 + kind, nsteps, dt, 1.0,
                                    sig, eps, rc, skin, T)
    assert np.isfinite(pe) and np.isfinite(temp)
    assert same, diff


def test_mc_barostat_bit_exact(oracle):
    """mc_barostat (integrators.f90:864-945) of the reference, driven through
    the harness with an all-pairs Lennard-Jones compute_potential and its
    random_number seeded, against the oracle's restatement with the same two
    numbers: accept, box, positions, forces and pot_eng bit for bit."""
    from oracle.oracle import Reference
    from tests._states import all_pairs, lattice_lj
    R = Reference(private=True)
    seen = set()
    for seed, (pres, dv) in enumerate([(0.5, 0.01), (50.0, 0.02), (0.01, 0.05), (3.0, 0.002), (50.0, 0.002),
                                       (1.0, 0.001), (0.5, 0.002), (3.0, 0.002)] * 2):
        rng = np.random.default_rng(200 + seed % 8)
        st = lattice_lj(rng, 108, 5.2)
        st["temp"] = 1.3
        R.load(st)
        u, acc = R.mc_barostat(seed + 1, pres, dv, 1.0, 1.0, 2.5)
        ref = st.copy()
        R.store(ref)
        box = np.zeros(3)
        R.lib.ref_get_box(box.ctypes.data_as(C.POINTER(C.c_double)))
        nl = all_pairs(st.n)

        def potential(s):
            oracle.call("or_initialize_forces", s)
            oracle.pair_loop(s, nl, LJ, [1.0, 1.0, 2.5])
            return s["pot_eng"]

        mine = st.copy()
        a = oracle.mc_barostat(mine, potential, pres, u, dv)
        assert a == acc, (seed, a, acc)
        seen.add(acc)
        assert np.array_equal(mine.box, box)
        for f in ("rx", "ry", "rz", "rx_old", "fx", "fz"):
            assert np.array_equal(mine[f], ref[f]), (seed, f)
        assert mine["pot_eng"] == ref["pot_eng"] and mine["density"] == ref["density"]
    assert seen == {True, False}


def test_dipole_reaction_field_bit_exact(oracle, reference):
    """dipole_calculate_forces, its side effects on the module's dipole and
    reaction-field arrays (the whole-array update of every particle's field,
    force.f90:2745-2747) and eval_reaction_field, bit for bit."""
    from tests._states import base_state
    rng = np.random.default_rng(77)
    n = 60
    st = base_state(rng, n, 6.0, charges=True)
    st["charge"] = st["charge"] * rng.uniform(0.5, 1.5, n)
    st["charge"][::7] = 0.0
    reference.load(st)
    reference.init_reaction_field()
    dip = np.zeros((3, n)); rfl = np.zeros((3, n))
    for t in range(600):
        i, j = rng.integers(0, n, 2)
        if i == j:
            continue
        a = oracle.dipole_pair(st, i, j, 2.4, 78.5, dip, rfl)
        b = reference.dipole_pair(i, j, 2.4, 78.5)
        assert a[0] == b[0] and all(np.array_equal(x, y) for x, y in zip(a[1:], b[1:])), t
    d, r = reference.get_dipole()
    assert np.array_equal(dip, d) and np.array_equal(rfl, r)
    assert np.count_nonzero(dip[0]) > n // 2
    ref = st.copy()
    reference.eval_reaction_field()
    reference.store(ref)
    mine = st.copy()
    oracle.eval_reaction_field(mine, dip, rfl)
    assert mine["pot_eng"] == ref["pot_eng"]
    for f in ("tx", "ty", "tz"):
        assert np.array_equal(mine[f], ref[f]), f


def test_composite_bodies_bit_exact(oracle, reference):
    """copy_transform (with and without the inertia bookkeeping) and
    transfer_forces (system.f90:743-881), bit for bit; satellites straddle
    the periodic boundary."""
    from tests._states import base_state
    rng = np.random.default_rng(5)
    n = 40
    st = base_state(rng, n, 5.0, sig=(1.0, 1.5, 2.0))
    st["mass"][20:] = 0.0
    st["ixx"][:5] = -1.0
    st["ixx"][25:30] = -1.0
    st["rx"][:10] = 2.45
    reference.load(st)
    mine = st.copy()
    ops = []
    for t in range(20):
        frm, to = t % 20, 20 + t
        off = rng.normal(size=3) * 0.8
        ops.append(("copy", frm, to, off, [None, False, True][t % 3]))
    for t in range(20):
        ops.append(("transfer", 20 + t, t % 7, None, None))
    for op, frm, to, off, si in ops:
        if op == "copy":
            assert oracle.copy_transform(mine, frm, to, off, si) == 0
            reference.copy_transform(frm, to, off, si)
        else:
            oracle.transfer_forces(mine, frm, to)
            reference.transfer_forces(frm, to)
    ref = st.copy()
    reference.store(ref)
    for f in ("rx", "ry", "rz", "qw", "qz", "ixx", "iyy", "izz", "fx", "fy", "fz", "tx", "ty", "tz"):
        assert np.array_equal(mine[f], ref[f]), f
