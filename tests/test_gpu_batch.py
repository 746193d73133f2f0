"""Batched replicas on the MI355X against the oracle run system by system.

A batched step advances every system through the README loop
(README.md:83-115): lgv_initial_step, initialize_forces, the double pair loop
over all pairs I < J of the system, lgv_final_step and calculate_thermo.  The
oracle runs each system as its own State with its own box, its pair loop
walking an all-pairs list in that same I < J order.  Tolerance as for the
single-system steps: 1e-10 of the field's RMS.
"""
import numpy as np
import pytest

from oracle.oracle import FIELDS, GB, LGV, LJ, Oracle
from tests._states import base_state, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def oracle():
    return Oracle()


def _spread(rng, n, L, dmin):
    """n points in a periodic cube of edge L, no two closer than dmin."""
    pts = []
    while len(pts) < n:
        p = rng.uniform(-L / 2, L / 2, 3)
        ok = True
        for q in pts:
            d = p - q
            d -= np.round(d / L) * L
            if d @ d < dmin * dmin:
                ok = False
                break
        if ok:
            pts.append(p)
    return np.array(pts)


# (kind, n, box, params, T)
SYSTEMS = [
    (LJ, 37, 5.0, (1.0, 1.0, 2.5), 1.0),
    (GB, 29, 13.0, (4.0, 2, 1), 3.0),
    (LJ, 70, 5.5, (1.1, 0.8, 2.5), 1.4),
    (GB, 45, 12.5, (3.5, 2, 1), 2.6),
    (LJ, 2, 4.0, (1.0, 1.0, 1.5), 0.7),
    (GB, 130, 14.0, (4.0, 2, 1), 3.3),
    (LJ, 300, 8.0, (1.0, 1.0, 2.5), 0.9),
]


def _systems(seed=4):
    from sucrose_amd import gen_ellipsoids_rsa
    rng = np.random.default_rng(seed)
    out = []
    for kind, n, L, p, T in SYSTEMS:
        if kind == GB:
            st = base_state(rng, n, L, sig=(1, 1, 3), eps=(5, 5, 1), T=T)
            pos, q = gen_ellipsoids_rsa(n, (0.5, 0.5, 1.5), L, int(rng.integers(1, 1 << 30)))
            assert len(pos) == n
            for k, f in enumerate(("qw", "qx", "qy", "qz")):
                st[f] = q[:, k]
        else:
            st = base_state(rng, n, L, T=T)
            pos = _spread(rng, n, L, 0.85)
        for k, f in enumerate(("rx", "ry", "rz")):
            st[f] = pos[:, k]
        out.append((kind, p, T, st))
    return out


def _load(dev, systems):
    buf = np.concatenate([st.buf for _, _, _, st in systems])
    dev.set_num_atoms(len(buf), 1)
    dev.set_box(max(st.box[0] for *_, st in systems))
    for k, f in enumerate(FIELDS):
        dev.upload(f, buf[:, k])
    dev.batch_set([st.n for *_, st in systems], [st.box[0] for *_, st in systems], [k for k, *_ in systems],
                  [p for _, p, _, _ in systems], [T for _, _, T, _ in systems])


def _all_pairs(n):
    return {"offset": np.concatenate([[0], np.cumsum(np.arange(n - 1, -1, -1))]).astype(np.int64),
            "nbr": np.concatenate([np.arange(i + 1, n) for i in range(n)] + [np.zeros(1)]).astype(np.int32),
            "x_last": np.zeros(n), "y_last": np.zeros(n), "z_last": np.zeros(n),
            "cutoff": 0.0, "skin": 0.0, "nnz": n * (n - 1) // 2}


def _oracle_steps(O, kind, p, st, nsteps, step0, dt, damping):
    st = st.copy()
    nl = _all_pairs(st.n)
    for t in range(nsteps):
        O.integrate(LGV, 0, st, dt, step=step0 + t, damping=damping)
        O.call("or_initialize_forces", st)
        O.pair_loop(st, nl, kind, list(p) + [0.0])
        O.integrate(LGV, 1, st, dt, damping=damping)
        O.call("or_calculate_thermo", st)
    return st


def _compare(dev, oracle, systems, nsteps, step0, dt, damping, label):
    th = dev.batch_thermo()
    fields = {f: dev.download(f.upper()) for f in ("rx", "ry", "rz", "vx", "vy", "vz", "qw", "qx", "qy", "qz",
                                                    "lx", "ly", "lz", "fx", "fy", "fz", "tx", "ty", "tz")}
    off = 0
    for s, (kind, p, T, st) in enumerate(systems):
        ref = _oracle_steps(oracle, kind, p, st, nsteps, step0, dt, damping)
        sl = slice(off, off + st.n)
        for f, v in fields.items():
            if kind == LJ and f in ("tx", "ty", "tz"):
                continue
            e = rel_err(v[sl], ref[f])
            assert e < TOL, (label, s, f, e)
        for c, name in enumerate(("kin_eng", "kin_rot", "pot_eng", "temp")):
            assert abs(th[s, c] - ref[name]) <= TOL * max(1.0, abs(ref[name])), (label, s, name, th[s, c], ref[name])
        off += st.n


def test_batched_steps_match_oracle(oracle):
    import sucrose_amd as sx
    systems = _systems()
    say(f"batched replicas: {len(systems)} systems, 3 steps vs oracle")
    dev = sx.System(0)
    _load(dev, systems)
    dev.batch_run(3, 0.002, damping=1.0, first_step=11)
    _compare(dev, oracle, systems, 3, 11, 0.002, 1.0, "graph")
    dev.close()


@pytest.mark.parametrize("shape", range(10))
def test_batched_launch_shapes(oracle, shape):
    import sucrose_amd as sx
    systems = _systems(seed=5 + shape)
    say(f"batched replicas: launch shape {sx.System.BATCH_SHAPES[shape]}, 1 step vs oracle")
    dev = sx.System(0)
    _load(dev, systems)
    dev.batch_set_shape(shape)
    dev.batch_run(1, 0.003, damping=0.5, first_step=4)
    _compare(dev, oracle, systems, 1, 4, 0.003, 0.5, f"shape {shape}")
    dev.close()


def test_batched_autotune_restores_state(oracle):
    import sucrose_amd as sx
    systems = _systems(seed=9)
    say("batched replicas: autotune, then 2 steps vs oracle")
    dev = sx.System(0)
    _load(dev, systems)
    choice, ms = dev.batch_autotune(0.002, damping=1.0, first_step=7, nrep=2)
    say(f"  chose {choice}: " + " ".join(f"{m:.3f}" for m in ms))
    assert 0 <= choice < len(sx.System.BATCH_SHAPES) and len(ms) == 10 and np.all(ms > 0)
    dev.batch_run(2, 0.002, damping=1.0, first_step=7)
    _compare(dev, oracle, systems, 2, 7, 0.002, 1.0, "autotuned")
    dev.close()


def test_batched_argument_errors():
    import sucrose_amd as sx
    systems = _systems(seed=12)
    say("batched replicas: argument errors")
    dev = sx.System(0)
    _load(dev, systems)
    counts = [st.n for *_, st in systems]
    boxes = [st.box[0] for *_, st in systems]
    kinds = [k for k, *_ in systems]
    params = [list(p) for _, p, _, _ in systems]
    temps = [T for _, _, T, _ in systems]
    bad = [list(q) for q in params]
    bad[0][2] = 0.5 * boxes[0] + 0.01          # LJ cutoff beyond half the box
    with pytest.raises(sx.SucroseError) as e:
        dev.batch_set(counts, boxes, kinds, bad, temps)
    assert e.value.code == -6   # SX_ERR_CUTOFF
    bad = [list(q) for q in params]
    bad[1][0] = 0.5 * boxes[1] - 1.99          # GB reach cutoff + sig_z - 1 beyond half the box
    with pytest.raises(sx.SucroseError) as e:
        dev.batch_set(counts, boxes, kinds, bad, temps)
    assert e.value.code == -6   # SX_ERR_CUTOFF
    with pytest.raises(sx.SucroseError):
        dev.batch_set(counts[:-1], boxes[:-1], kinds[:-1], params[:-1], temps[:-1])
    dev.close()
