"""FP32 pair mode (sx_set_pair_precision(32)) vs the CPU oracle.

The workloads quote FP32 pair terms.  In that mode the device evaluates the
LJ, WCA and Debye-Hueckel pair terms in FP32 (hardware reciprocal, square
root and exp2) and gb_ (variant 0) in a form with fewer FP64 divisions (FP32
is too coarse for gb_ near contact), while positions, separations, the cutoff
and neighbour decisions, every sum over pairs and the integrators stay FP64.
The bar: every force and torque component within 1e-5 of its own magnitude
plus the median magnitude, and the potential energy within 1e-5 relative,
against the oracle's FP64 reference arithmetic.  Each case also runs the
same state at 64 bits, whose error is stated beside the FP32 one and held
to 1e-10; the FP32 error must exceed 1e-9, which shows the FP32 path ran.
"""
import numpy as np
import pytest

from oracle.oracle import DH, GB, LGVC, LJ, WCA
from tests._pairs import close
from tests._states import base_state, from_device, rel_err, to_device
from tests.test_gpu_parity import overlap_free_state

pytestmark = pytest.mark.gpu

F32_TOL = 1e-5
F64_TOL = 1e-10


def say(*a):
    print(*a, flush=True)


F32_CASES = [
    (LJ, (1.0, 1.0, 2.5), 2.5, dict(), 2000, 13.0),
    (WCA, (1.0, 1.0), 1.2, dict(), 2000, 13.0),
    (DH, (5.0, 1.0, 1.0), 5.0, dict(charges=True), 2000, 16.0),
    (GB, (4.0, 2, 1), 6.0, dict(sig=(1, 1, 3), eps=(5, 5, 1)), 1500, 22.0),
    (GB, (4.0, 1, 2), 6.0, dict(sig=(1, 1, 3), eps=(5, 5, 1)), 1500, 22.0),
]


@pytest.mark.parametrize("kind,params,rc,kw,n,L", F32_CASES)
def test_f32_forces_vs_oracle(oracle, kind, params, rc, kw, n, L):
    """One pair loop in each precision vs the oracle's FP64 loop."""
    import sucrose_amd as sx
    say(f"f32 forces kind={kind} params={params} n={n}")
    rng = np.random.default_rng(11 * kind + n)
    st = overlap_free_state(rng, kind, n, L, kw, seed=11 * kind + n)
    nl = oracle.nlist_new(st, rc, 0.3)
    oracle.nlist_build(st, nl)
    ref = st.copy()
    oracle.pair_loop(ref, nl, kind, list(params))
    comps = ("fx", "fy", "fz", "tx", "ty", "tz") if kind == GB else ("fx", "fy", "fz")
    assert np.count_nonzero(ref["fx"]) > n // 4
    dev = sx.System(0)
    errs = {}
    for bits in (64, 32):
        to_device(dev, st)
        dev.set_pair_precision(bits)
        assert dev.pair_precision() == bits
        dev.init_nlist(0.3, rc)
        dev.initialize_forces()
        dev.forces(kind, *params)
        got = from_device(dev, st)
        e = max(close(got[f], ref[f]) for f in comps)
        epe = abs(got["pot_eng"] - ref["pot_eng"]) / abs(ref["pot_eng"])
        errs[bits] = (e, epe)
        say(f"  {bits}-bit pair terms: worst element error {e:.3e}, potential energy {epe:.3e}")
    dev.close()
    assert errs[64][0] <= F64_TOL and errs[64][1] <= F64_TOL, errs
    assert errs[32][0] <= F32_TOL and errs[32][1] <= F32_TOL, errs
    assert errs[32][0] > 1e-9, errs


def test_f32_precision_argument():
    import sucrose_amd as sx
    dev = sx.System(0)
    assert dev.pair_precision() == 64
    with pytest.raises(sx.SucroseError):
        dev.set_pair_precision(16)
    dev.set_pair_precision(32)
    assert dev.pair_precision() == 32
    dev.close()


def test_f32_charged_chain_step(oracle):
    """One charged-chain step of the step driver (LGVC + RATTLE, the WCA and
    Debye-Hueckel terms in one pass, bonded terms) with FP32 pair terms vs
    the oracle's FP64 step: forces to 1e-5, positions and velocities to 1e-7
    of their RMS (the FP32 force error enters them through dt)."""
    import sucrose_amd as sx
    from tests.test_gpu_rows import _bonded_terms, _oracle_bonded
    say("f32 charged chains: 30 x 50 beads, one step")
    nchains, nbeads = 30, 50
    n = nchains * nbeads
    L = (n / 0.5) ** (1.0 / 3.0)
    rng = np.random.default_rng(4321)
    st = base_state(rng, n, L, forces=False, T=1.0, charges=True)
    st["charge"] = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    pos = sx.gen_chains(nchains, nbeads, L, bond=1.0, min_dist=0.95, seed=12)
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
    dev.init_nlist(0.8, 5.0)
    dev.initialize_forces()
    for kind, params in loops:
        dev.forces(kind, *params)
    dev.bonded_forces()
    start = from_device(dev, st)
    start.bbi, start.bbj, start.bb_len = st.bbi, st.bbj, st.bb_len
    ref = start.copy()
    nl = oracle.nlist_new(ref, 5.0, 0.8)
    oracle.nlist_build(ref, nl)
    dt, first = 0.005, 5
    oracle.integrate(LGVC, 0, ref, dt, step=first, damping=1.0)
    oracle.call("or_initialize_forces", ref)
    for kind, params in loops:
        oracle.pair_loop(ref, nl, kind, list(params))
    _oracle_bonded(oracle, ref, terms)
    oracle.integrate(LGVC, 1, ref, dt, damping=1.0)
    oracle.call("or_calculate_thermo", ref)
    dev.set_pair_precision(32)
    dev.run(sx.INT_LGVC, [(k,) + p for k, p in loops], 1, dt, damping=1.0, first_step=first, bonded=True)
    got = from_device(dev, st)
    dev.close()
    ef = max(close(got[f], ref[f]) for f in ("fx", "fy", "fz"))
    ex = max(rel_err(got[f], ref[f]) for f in ("rx", "ry", "rz", "vx", "vy", "vz"))
    epe = abs(got["pot_eng"] - ref["pot_eng"]) / abs(ref["pot_eng"])
    say(f"  FP32 pair terms: forces {ef:.3e}, positions/velocities {ex:.3e}, potential energy {epe:.3e}")
    assert ef <= F32_TOL and epe <= F32_TOL
    assert ex <= 1e-7
    assert ef > 1e-9


def test_f32_batched_step(oracle):
    """Batched LJ and GB(3,5,2,1) replicas, one step with FP32 pair terms in
    every launch shape vs the oracle's FP64 systems: forces, torques and
    energies to 1e-5, positions and velocities to 1e-5 of their RMS."""
    import sucrose_amd as sx
    from tests.test_gpu_batch import _load, _oracle_steps, _systems
    systems = _systems(seed=21)
    refs = [_oracle_steps(oracle, kind, p, st, 1, 4, 0.003, 0.5) for kind, p, T, st in systems]
    worst = 0.0
    for shape in range(len(sx.System.BATCH_SHAPES)):
        say(f"f32 batched replicas: launch shape {sx.System.BATCH_SHAPES[shape]}, 1 step vs oracle")
        dev = sx.System(0)
        _load(dev, systems)
        dev.set_pair_precision(32)
        dev.batch_set_shape(shape)
        dev.batch_run(1, 0.003, damping=0.5, first_step=4)
        th = dev.batch_thermo()
        fields = {f: dev.download(f.upper()) for f in ("rx", "ry", "rz", "vx", "vy", "vz", "fx", "fy", "fz",
                                                        "tx", "ty", "tz")}
        dev.close()
        off = 0
        for s, ((kind, p, T, st), ref) in enumerate(zip(systems, refs)):
            sl = slice(off, off + st.n)
            comps = ("fx", "fy", "fz", "tx", "ty", "tz") if kind == GB else ("fx", "fy", "fz")
            ef = max(close(fields[f][sl], ref[f]) for f in comps)
            ex = max(rel_err(fields[f][sl], ref[f]) for f in ("rx", "ry", "rz", "vx", "vy", "vz"))
            epe = abs(th[s, 2] - ref["pot_eng"]) / max(1.0, abs(ref["pot_eng"]))
            assert ef <= F32_TOL and epe <= F32_TOL, (shape, s, ef, epe)
            assert ex <= F32_TOL, (shape, s, ex)
            worst = max(worst, ef)
            off += st.n
    say(f"  worst FP32 force element error over shapes and systems: {worst:.3e}")
    assert worst > 1e-9
