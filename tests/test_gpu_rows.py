"""Device vs oracle for the charged-chain step, the constrained Nose-Hoover
step, thermo's virial and pressure and the velocity initialisation, and the
step driver's contracts: a new target temperature, continuing the Langevin
counter, and a neighbour list that outgrows its storage inside the replays.

Tolerances as in test_gpu_parity.py: FP64 in the reference's order on both
sides, summation order and libm ulps aside.
"""
import ctypes as C

import numpy as np
import pytest

from oracle.oracle import DH, LGV, LGVC, LJ, MORSE, NHTC, VV, WCA
from tests._pairs import close
from tests._states import base_state, from_device, lattice_lj, rel_err, to_device

pytestmark = pytest.mark.gpu
TOL = 1e-10


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dev():
    import sucrose_amd as sx
    s = sx.System(0)
    yield s
    s.close()


# --------------------------------------------------------------------------
# the charged-chain step (LGVC + RATTLE + WCA + Debye-Hueckel + bonded terms)
# --------------------------------------------------------------------------
def _bonded_terms(nchains, nbeads):
    ids = np.arange(nchains * nbeads).reshape(nchains, nbeads)
    bends = np.stack([ids[:, :-2], ids[:, 1:-1], ids[:, 2:]], -1).reshape(-1, 3)
    dihs = np.stack([ids[:, :-3], ids[:, 1:-2], ids[:, 2:-1], ids[:, 3:]], -1).reshape(-1, 4)
    # cosine_bend_ about the straight chain (force.f90:2345) and dih_ (:2286)
    return [(2, bends, 20.0, 180.0), (3, dihs, 1.0, 0.0)]


def _oracle_bonded(O, st, terms):
    """The caller's loop over the bonded routines: forces onto the atoms,
    energies onto pot_eng."""
    for kind, idx, coeff, angle in terms:
        for t in idx:
            e, f = O.bonded(kind, st, list(t), coeff, angle)
            st["pot_eng"] = st["pot_eng"] + e
            for p, a in enumerate(t):
                st["fx"][a] += f[p][0]
                st["fy"][a] += f[p][1]
                st["fz"][a] += f[p][2]


def test_run_graph_charged_chains(oracle):
    """sx_run(LGVC, [WCA, DH], bonded) -- the charged-chain step -- on 40
    self-avoiding chains of 50 beads at bead density 0.5 (RATTLE bonds of
    length 1, alternating charges), 12 graph-replayed steps with a skin small
    enough that the list is rebuilt on the way, vs the oracle's README loop
    (lgvc_initial_step with apply_constraints_a, the rebuild check, the WCA
    and Debye-Hueckel pair loops, the bonded terms, lgvc_final_step with
    apply_constraints_b, calculate_thermo)."""
    import sucrose_amd as sx
    say("run graph charged chains: 40 x 50 beads, LGVC + RATTLE + WCA + DH + bonded")
    nchains, nbeads = 40, 50
    n = nchains * nbeads
    L = (n / 0.5) ** (1.0 / 3.0)
    rng = np.random.default_rng(1234)
    st = base_state(rng, n, L, forces=False, T=1.0, charges=True)
    st["charge"] = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    pos = sx.gen_chains(nchains, nbeads, L, bond=1.0, min_dist=0.95, seed=11)
    st["rx"], st["ry"], st["rz"] = pos[:, 0], pos[:, 1], pos[:, 2]
    ids = np.arange(n).reshape(nchains, nbeads)
    st.bbi = ids[:, :-1].ravel().astype(np.int32)
    st.bbj = ids[:, 1:].ravel().astype(np.int32)
    st.bb_len = np.ones(len(st.bbi))
    st["dof"] = 6 * n - st.nc - 3
    terms = _bonded_terms(nchains, nbeads)
    loops = [(WCA, (1.0, 1.0)), (DH, (5.0, 1.0, 1.0))]
    dev = sx.System(0)
    to_device(dev, st)
    for kind, idx, coeff, angle in terms:
        dev.set_bonded(kind, idx + 1, coeff, angle)
    skin = 0.04
    dev.init_nlist(skin, 5.0)
    dev.set_rebuild_rule(sx.REBUILD_HALF_SKIN)
    dev.initialize_forces()
    for kind, params in loops:
        dev.forces(kind, *params)
    dev.bonded_forces()
    ref = from_device(dev, st)
    ref.bbi, ref.bbj, ref.bb_len = st.bbi, st.bbj, st.bb_len
    nl = oracle.nlist_new(ref, 5.0, skin)
    oracle.nlist_build(ref, nl)
    nsteps, dt, first = 12, 0.005, 3
    dev.run(sx.INT_LGVC, [(k,) + p for k, p in loops], nsteps, dt, damping=1.0, first_step=first, bonded=True)
    rebuilds = 0
    for t in range(nsteps):
        oracle.integrate(LGVC, 0, ref, dt, step=first + t, damping=1.0)
        d = np.stack([ref[f] - nl[x] for f, x in (("rx", "x_last"), ("ry", "y_last"), ("rz", "z_last"))], 1)
        d -= np.round(d / L) * L
        if (d * d).sum(1).max() > (skin / 2) ** 2:
            oracle.nlist_build(ref, nl)
            rebuilds += 1
        oracle.call("or_initialize_forces", ref)
        for kind, params in loops:
            oracle.pair_loop(ref, nl, kind, list(params))
        _oracle_bonded(oracle, ref, terms)
        oracle.integrate(LGVC, 1, ref, dt, damping=1.0)
        oracle.call("or_calculate_thermo", ref)
    got = from_device(dev, st)
    dev.close()
    dr = np.stack([got[f][st.bbi] - got[f][st.bbj] for f in ("rx", "ry", "rz")], 1)
    dr -= np.round(dr / L) * L
    bl = np.sqrt((dr * dr).sum(1))
    say(f"  {rebuilds} rebuilds; bond lengths {bl.min():.6f}..{bl.max():.6f}")
    assert rebuilds >= 1
    assert np.all(np.abs(bl - 1.0) < 1e-4)
    for f in ("rx", "ry", "rz", "vx", "vy", "vz", "fx", "fy", "fz", "qw", "lx"):
        e = rel_err(got[f], ref[f])
        assert e < 1e-9, (f, e)
    for s in ("pot_eng", "kin_eng", "kin_rot", "temp", "virial", "virial_bonded", "pressure"):
        assert abs(got[s] - ref[s]) <= 1e-9 * max(1.0, abs(ref[s])), (s, got[s], ref[s])


# --------------------------------------------------------------------------
# nhtc_initial_step / nhtc_final_step (integrators.f90:394-490)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("nchains", [16, 60])
@pytest.mark.parametrize("phase", [0, 1])
def test_nhtc_one_step(dev, oracle, nchains, phase):
    """One constrained Nose-Hoover phase (the chain update, then RATTLE a or
    b) on chains of 40 beads: 640 particles walk the chain recurrence
    serially, 2400 take the parallel evaluation."""
    from tests._states import chains
    say(f"nhtc phase={phase} chains={nchains}")
    rng = np.random.default_rng(500 + nchains + phase)
    n = nchains * 40
    Lb = 9.0 if nchains == 16 else 14.0
    st = base_state(rng, n, Lb, sig=(1, 1, 3))
    pos, bi, bj = chains(rng, nchains, 40, np.full(3, Lb))
    st["rx"], st["ry"], st["rz"] = pos[:, 0], pos[:, 1], pos[:, 2]
    st.bbi, st.bbj, st.bb_len = bi, bj, np.ones(len(bi))
    st["dof"] = 6 * n - len(bi) - 3
    # near the chain's thermal regime: sum m v^2 at n_dof T, weak forces
    s2 = np.sum(st["mass"] * (st["vx"] ** 2 + st["vy"] ** 2 + st["vz"] ** 2))
    f = np.sqrt((3 * n - 3) * 1.5 / s2)
    for c in ("vx", "vy", "vz"):
        st[c] = st[c] * f
    for c in ("fx", "fy", "fz"):
        st[c] = st[c] * 0.1
    st["peta0"], st["peta1"], st["peta2"] = 0.3, -0.2, 0.1
    to_device(dev, st)
    ref = st.copy()
    dt = 0.0015
    oracle.integrate(NHTC, phase, ref, dt)
    (dev.nhtc_initial_step if phase == 0 else dev.nhtc_final_step)(dt)
    got = from_device(dev, st)
    for fl in ("rx", "ry", "rz", "vx", "vy", "vz"):
        assert rel_err(got[fl], ref[fl]) < TOL, fl
    for s in ("peta0", "peta1", "peta2", "eta0", "eta1", "eta2"):
        assert abs(got[s] - ref[s]) <= 1e-8 * max(1.0, abs(ref[s])), (s, got[s], ref[s])
    assert abs(got["virial_bonded"] - ref["virial_bonded"]) <= TOL * max(1.0, abs(ref["virial_bonded"]))
    if phase == 1:
        assert abs(got["kin_eng"] - ref["kin_eng"]) <= TOL * abs(ref["kin_eng"])


# --------------------------------------------------------------------------
# calculate_thermo's virial and pressure (system.f90:725-741)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("constrained", [False, True])
def test_thermo_virial_pressure(dev, oracle, constrained):
    """Temperature, virial (sum r.F over the wrapped positions, plus the
    bonded virial when there are constraints) and pressure after the LJ pair
    loop, against the oracle on the same state and forces."""
    say(f"thermo virial/pressure constrained={constrained}")
    rng = np.random.default_rng(71 + constrained)
    n, L = 1000, 11.0
    st = lattice_lj(rng, n, L, jitter=0.05)
    # some positions outside the box: the virial uses the wrapped ones
    st["rx"][::7] += L
    st["ry"][::11] -= L
    if constrained:
        st.bbi = np.arange(0, 200, 2, dtype=np.int32)
        st.bbj = st.bbi + 1
        st.bb_len = np.ones(len(st.bbi))
        st["virial_bonded"] = 3.25
    st["dof"] = 6 * n - st.nc - 3
    st["kin_eng"], st["kin_rot"] = 1234.5, 432.1
    to_device(dev, st)
    dev.init_nlist(0.3, 2.5)
    dev.initialize_forces()
    dev.forces(LJ, 1.0, 1.0, 2.5)
    dev.set_scalar("KIN_ENG", st["kin_eng"])
    dev.set_scalar("KIN_ROT", st["kin_rot"])
    dev.set_scalar("VIRIAL_BONDED", st["virial_bonded"])
    ref = from_device(dev, st)     # the device's forces, then thermo on both sides
    ref.bbi, ref.bbj, ref.bb_len = st.bbi, st.bbj, st.bb_len
    dev.calculate_thermo()
    oracle.call("or_calculate_thermo", ref)
    got = from_device(dev, st)
    say(f"  virial {got['virial']:.10g} vs {ref['virial']:.10g}, pressure {got['pressure']:.10g} vs "
        f"{ref['pressure']:.10g}")
    assert ref["virial"] != 0 and ref["pressure"] != 0
    for s in ("temp", "virial", "pressure"):
        assert abs(got[s] - ref[s]) <= 1e-12 * abs(ref[s]), (s, got[s], ref[s])


# --------------------------------------------------------------------------
# set_velocity_to_temperature (system.f90:362-421)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("remove_com", [1, 0])
def test_set_velocity_to_temperature(dev, oracle, remove_com):
    """Philox normals from the (atom, int(1000 T), seed, 42) counters, the
    mean velocity removed, and the scaling to sum v^2 = dof T Nrigid,
    against the oracle."""
    say(f"set_velocity_to_temperature remove_com={remove_com}")
    rng = np.random.default_rng(80 + remove_com)
    n = 3000
    st = base_state(rng, n, 20.0)
    st.bbi = np.arange(0, 100, 2, dtype=np.int32)
    st.bbj = st.bbi + 1
    st.bb_len = np.ones(len(st.bbi))
    to_device(dev, st)
    T, seed = 1.785, 15
    dev.set_velocity_to_temperature(T, seed, remove_com=bool(remove_com))
    got = from_device(dev, st)
    ref = st.copy()
    oracle.call("or_set_velocity_to_temperature", ref, C.c_double(T), C.c_int(seed), C.c_int(remove_com))
    for f in ("vx", "vy", "vz"):
        assert rel_err(got[f], ref[f]) < 1e-12, f
    # the reference removes the mean velocity and scales sum v^2 (unweighted)
    p = [np.sum(got[f]) for f in ("vx", "vy", "vz")]
    if remove_com:
        assert max(abs(x) for x in p) < 1e-10 * np.sqrt(n)
    s2 = np.sum(got["vx"] ** 2 + got["vy"] ** 2 + got["vz"] ** 2)
    dof = 6 * n - st.nc - (3 if remove_com else 0)
    assert abs(s2 - dof * T) <= 1e-12 * dof * T


# --------------------------------------------------------------------------
# the step driver's contracts
# --------------------------------------------------------------------------
def _oracle_lgv(O, ref, nl, kind, p, nsteps, first, dt):
    for t in range(nsteps):
        O.integrate(LGV, 0, ref, dt, step=first + t, damping=1.0)
        if O.needs_rebuild(ref, nl):
            O.nlist_build(ref, nl)
        O.call("or_initialize_forces", ref)
        O.pair_loop(ref, nl, kind, p)
        O.integrate(LGV, 1, ref, dt, damping=1.0)
        O.call("or_calculate_thermo", ref)


def _lj_langevin_start(dev, oracle, seed):
    rng = np.random.default_rng(seed)
    st = lattice_lj(rng, 864, 10.0776, T=1.0)
    to_device(dev, st)
    dev.init_nlist(0.3, 2.5)
    dev.initialize_forces()
    dev.forces(LJ, 1.0, 1.0, 2.5)
    ref = from_device(dev, st)
    nl = oracle.nlist_new(ref, 2.5, 0.3)
    oracle.nlist_build(ref, nl)
    return st, ref, nl


def _same(got, ref, tol=1e-9):
    for f in ("rx", "vx", "vz", "qw", "lx", "fx"):
        assert rel_err(got[f], ref[f]) < tol, (f, rel_err(got[f], ref[f]))
    for s in ("pot_eng", "kin_eng", "kin_rot", "temp"):
        assert abs(got[s] - ref[s]) <= tol * max(1.0, abs(ref[s])), (s, got[s], ref[s])


def test_run_new_target_temperature(dev, oracle):
    """Langevin steps, a new target temperature, then more steps with the
    same step spec: the second run uses the new temperature and its noise
    counter int(1000 T) (integrators.f90:532, :546-547), as the oracle's
    loop does."""
    import sucrose_amd as sx
    say("run graph: Langevin, new target temperature, Langevin again")
    st, ref, nl = _lj_langevin_start(dev, oracle, 41)
    spec = [(LJ, 1.0, 1.0, 2.5)]
    dev.run(sx.INT_LGV, spec, 6, 0.004, first_step=1)
    dev.set_target_temperature(2.5)
    dev.run(sx.INT_LGV, spec, 6, 0.004, first_step=7)
    got = from_device(dev, st)
    _oracle_lgv(oracle, ref, nl, LJ, [1.0, 1.0, 2.5], 6, 1, 0.004)
    ref["target_temp"] = 2.5
    _oracle_lgv(oracle, ref, nl, LJ, [1.0, 1.0, 2.5], 6, 7, 0.004)
    _same(got, ref)


def test_run_continues_langevin_counter(dev, oracle):
    """first_step < 0 continues the device's Langevin counter: 5 steps from
    step 9, then 5 more without a step number, equal the oracle's steps
    9..18."""
    import sucrose_amd as sx
    say("run graph: Langevin counter continued with first_step = -1")
    st, ref, nl = _lj_langevin_start(dev, oracle, 43)
    spec = [(LJ, 1.0, 1.0, 2.5)]
    dev.run(sx.INT_LGV, spec, 5, 0.004, first_step=9)
    dev.run(sx.INT_LGV, spec, 5, 0.004, first_step=-1)
    got = from_device(dev, st)
    _oracle_lgv(oracle, ref, nl, LJ, [1.0, 1.0, 2.5], 10, 9, 0.004)
    _same(got, ref)


def test_run_list_outgrows_storage(dev, oracle):
    """A rebuild inside sx_run that needs more list storage than the list
    has: the call reports SX_ERR_CAPACITY, the list is grown and rebuilt at
    the current positions (equal to the oracle's list there), and the next
    call steps on as the oracle's loop does."""
    import sucrose_amd as sx
    say("run: list outgrows its storage inside the step (positions squeezed to 0.55)")
    rng = np.random.default_rng(45)
    n, L = 864, 10.0776
    st = lattice_lj(rng, n, L, T=1.0)
    to_device(dev, st)
    morse = (1.0, 2.0, 1.1, 2.5)
    dev.init_nlist(0.3, 2.5)
    dev.initialize_forces()
    dev.forces(MORSE, *morse)
    # the same particles squeezed into the middle 17% of the box: about 3.7x
    # the entries, more than the storage sized for the lattice (2.1x) holds
    for f in ("RX", "RY", "RZ"):
        dev.upload(f, dev.download(f) * 0.55)
    with pytest.raises(sx.SucroseError) as e:
        dev.run(sx.INT_VV, [(MORSE,) + morse], 4, 0.002)
    say(f"  {e.value}")
    assert e.value.code == -4
    ref = from_device(dev, st)
    off, nb = dev.export_nlist()
    nl = oracle.nlist_new(ref, 2.5, 0.3)
    oracle.nlist_build(ref, nl)
    say(f"  list after recovery: {len(nb)} pairs (oracle {nl['nnz']})")
    assert len(nb) == nl["nnz"] and np.array_equal(nb - 1, nl["nbr"][:nl["nnz"]])
    assert np.array_equal(off[:-1] - 1, nl["offset"][:n - 1]) and off[-1] == nl["nnz"] + 1
    # the run goes on from the recovered list
    dev.initialize_forces()
    dev.forces(MORSE, *morse)
    ref = from_device(dev, st)
    dev.run(sx.INT_VV, [(MORSE,) + morse], 6, 0.002)
    for t in range(6):
        oracle.integrate(VV, 0, ref, 0.002)
        if oracle.needs_rebuild(ref, nl):
            oracle.nlist_build(ref, nl)
        oracle.call("or_initialize_forces", ref)
        oracle.pair_loop(ref, nl, MORSE, list(morse))
        oracle.integrate(VV, 1, ref, 0.002)
        oracle.call("or_calculate_thermo", ref)
    got = from_device(dev, st)
    for f in ("rx", "vx", "fx", "fz"):
        assert close(got[f], ref[f]) < 1e-9, (f, close(got[f], ref[f]))
    assert abs(got["pot_eng"] - ref["pot_eng"]) <= 1e-9 * abs(ref["pot_eng"])
