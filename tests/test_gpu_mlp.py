"""The lambda* warm-start network on the MI355X against its CPU oracle
(oracle/mlp.py), the training-sample pipeline against the oracle's contact
solver, and the learned ECP solve against the plain converged one.

Tolerances.  The device multiplies BF16 operands with FP32 accumulation; the
oracle rounds the same operands to BF16 and accumulates in FP64.  They differ
by FP32 accumulation (~1e-7 relative) and by the rare operand whose FP32 and
FP64 values round to neighbouring BF16 numbers (one BF16 step, 2^-8 relative,
on that operand only).  Hence: outputs within 2e-3 absolute at worst and
1e-4 on average; the gradient within 1e-2 of its norm (the loss within 1e-4
relative).  Contact parameters of the two solvers agree to 1e-9 and
converged-mode forces to 1e-10 of the oracle's.
"""
import numpy as np
import pytest

from oracle import mlp
from oracle.oracle import ECPR, Oracle
from tests._states import base_state, rand_quats, to_device

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dev():
    import sucrose_amd as sx
    s = sx.System(0)
    s.mlp_init(seed=11)
    yield s
    s.close()


def _inputs(rng, n):
    X = np.zeros((n, 7))
    r = rng.normal(size=(n, 3))
    X[:, :3] = r / np.linalg.norm(r, axis=1)[:, None]
    q = rand_quats(rng, n)
    q *= np.sign(q[:, :1])
    X[:, 3:] = q
    return X.astype(np.float32)


def test_forward_matches_oracle(dev):
    say("mlp forward: 4096 inputs vs oracle")
    rng = np.random.default_rng(0)
    p = (rng.normal(size=mlp.NPARAM) * 0.25).astype(np.float32)
    dev.mlp_set_params(p)
    X = _inputs(rng, 4096)
    Y = dev.mlp_forward(X)
    ref = mlp.forward(p.astype(np.float64), X.astype(np.float64), bf=True)
    err = np.abs(Y - ref)
    say(f"  max {err.max():.2e} mean {err.mean():.2e}")
    assert err.max() < 2e-3 and err.mean() < 1e-4


def test_gradient_matches_oracle(dev):
    say("mlp loss gradient: batch 256 vs oracle")
    rng = np.random.default_rng(1)
    p = (rng.normal(size=mlp.NPARAM) * 0.25).astype(np.float32)
    dev.mlp_set_params(p)
    X = _inputs(rng, 256)
    T = np.stack([rng.uniform(0.25, 0.75, 256), rng.uniform(0.33, 1.0, 256)], 1).astype(np.float32)
    loss, g = dev.mlp_loss_grad(X, T, (1.0, 0.5))
    lref, gref = mlp.loss_grad(p.astype(np.float64), X.astype(np.float64), T.astype(np.float64), (1.0, 0.5),
                               bf=True)
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    say(f"  loss {loss:.6e} vs {lref:.6e}; gradient rel err {rel:.2e}")
    assert abs(loss - lref) <= 1e-4 * abs(lref)
    assert rel < 1e-2
    # each layer's weight gradient on its own
    k = 0
    for o, i in mlp.SIZES:
        sl = slice(k, k + o * i)
        assert np.linalg.norm(g[sl] - gref[sl]) <= 2e-2 * np.linalg.norm(gref[sl]) + 1e-9
        k += o * i + o


def _ecp_state(n=3000, seed=5):
    import sucrose_amd as sx
    from sucrose_amd.workloads import ecp_langevin_state
    d = sx.System(0)
    p = ecp_langevin_state(d, n=n, seed=seed, stages=6, steps_per_stage=20, equil_steps=40, log=lambda *a: None)
    return d, p


def test_training_samples_match_oracle_contact_solve(dev):
    """Each sample's targets follow from its own inputs: in particle i's body
    frame A = S^2, B = R(q_rel) S^2 R(q_rel)^T, r = rhat, so the oracle's
    converged lambda* and contact distance reproduce them."""
    say("mlp samples: ECP 3000 state, targets vs oracle lambda*")
    d, p = _ecp_state()
    d.mlp_init(seed=3)
    tot = d.mlp_collect(ECPR, capacity=1 << 20)
    say(f"  {tot} samples")
    assert tot > 1000
    X, T = d.mlp_samples(0, min(tot, 400))
    O = Oracle()
    S2 = np.diag([0.25, 0.25, 2.25])
    for x, t in zip(X.astype(np.float64), T.astype(np.float64)):
        w, qx, qy, qz = x[3:] / np.linalg.norm(x[3:])
        R = np.array([[1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - w * qz), 2 * (qx * qz + w * qy)],
                      [2 * (qx * qy + w * qz), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - w * qx)],
                      [2 * (qx * qz - w * qy), 2 * (qy * qz + w * qx), 1 - 2 * (qx * qx + qy * qy)]])
        B = R @ S2 @ R.T
        lam, _ = O.lambda_converged(S2.ravel(order="F"), B.ravel(order="F"), x[:3])
        sig = O.contact_distance(S2.ravel(order="F"), B.ravel(order="F"), x[:3], lam)
        assert abs(lam - t[0]) < 2e-6, (lam, t)
        assert abs(sig / 3.0 - t[1]) < 2e-6, (sig, t)
    d.close()


def test_converged_solver_matches_oracle(dev):
    say("ECPR converged lambda: pair_eval vs oracle (400 particles)")
    import sucrose_amd as sx
    rng = np.random.default_rng(9)
    st = base_state(rng, 400, 9.0, sig=(1, 1, 3))
    d = sx.System(0)
    to_device(d, st)
    d.init_nlist(0.3, 3.1)
    d.set_ecp_solver(sx.ECP_CONVERGED)
    i = rng.integers(0, 400, 3000)
    j = rng.integers(0, 400, 3000)
    m = i != j
    i, j = i[m], j[m]
    pe, f, ti, tj, lam = d.pair_eval(ECPR, i + 1, j + 1)
    O = Oracle()
    nz = 0
    for t in range(len(i)):
        ar, br, r = O.ecp_shape(st, i[t], j[t])
        lr, _ = O.lambda_converged(ar, br, r)
        assert abs(lam[t] - lr) < 1e-12
        o = O.ecp_at_lambda(ECPR, st, i[t], j[t], [0.0], lr)
        nz += o[0] != 0
        for a, b in ((pe[t], o[0]), (f[t], o[1]), (ti[t], o[2]), (tj[t], o[3])):
            assert np.abs(np.asarray(a) - np.asarray(b)).max() <= 1e-10 * max(1.0, np.abs(b).max())
    assert nz > 20
    d.close()


def test_train_learned_solver(dev):
    """Training on solver-generated pairs lowers the loss; the learned warm
    start gives the plain converged solver's lambda and forces with fewer
    Newton iterations."""
    import sucrose_amd as sx
    say("mlp training on ECP samples, learned vs plain converged solve")
    d, p = _ecp_state(n=4000, seed=8)
    d.mlp_init(seed=21)
    for k in range(6):
        d.run(sx.INT_LGV, [(sx.ECPR,)], 20, p["dt"], damping=p["gamma"], first_step=p["next_step"] + 20 * k)
        tot = d.mlp_collect(ECPR, capacity=1 << 20)
    say(f"  {tot} samples")
    t1 = d.mlp_train(3000, batch=2048, lr=3e-3, w=(1.0, 0.1), seed=1)
    t2 = d.mlp_train(2000, batch=2048, lr=1e-3, w=(1.0, 0.1), seed=2)
    first, last = float(np.mean(t1[:20])), float(np.mean(t2[-50:]))
    say(f"  loss {first:.3e} -> {last:.3e}")
    assert last < 0.05 * first
    cmp = d.compare_solvers(ECPR, nrep=3)
    say("  " + " ".join(f"{k}={v:.4g}" for k, v in cmp.items()))
    assert cmp["pairs"] > 1000
    assert cmp["lambda_warm_vs_plain_max"] < 1e-9
    assert cmp["iters_warm"] < cmp["iters_plain"]
    assert cmp["lambda0_err_rms"] < 0.06   # 30k samples; the 32k-particle benchmark trains on 2M
    assert cmp["proven_out_of_range"] > 0
    # forces and torques of the learned solve equal the plain converged ones
    out = {}
    for solver in (sx.ECP_CONVERGED, sx.ECP_LEARNED):
        d.set_ecp_solver(solver)
        d.initialize_forces()
        d.forces(sx.ECPR)
        out[solver] = np.stack([d.download(f) for f in ("FX", "FY", "FZ", "TX", "TY", "TZ")])
    a, b = out[sx.ECP_CONVERGED], out[sx.ECP_LEARNED]
    rel = np.abs(a - b).max() / np.abs(a).max()
    say(f"  learned vs plain forces/torques: max rel {rel:.2e}")
    assert rel < 1e-9
    d.close()
