"""Pins of the network oracle (oracle/mlp.py): its FP64 gradient equals
torch's autograd of the same network and central differences of its loss; its
BF16 rounding matches torch's bfloat16 conversion; one Adam step matches
torch.optim.Adam."""
import numpy as np
import torch

from oracle import mlp


def _torch_net(p):
    layers = mlp.unpack(p)
    mods = []
    for k, (W, b) in enumerate(layers):
        lin = torch.nn.Linear(W.shape[1], W.shape[0]).double()
        with torch.no_grad():
            lin.weight.copy_(torch.tensor(W))
            lin.bias.copy_(torch.tensor(b))
        mods.append(lin)
        if k < len(layers) - 1:
            mods.append(torch.nn.SiLU())
    return torch.nn.Sequential(*mods)


def _batch(rng, n):
    X = rng.normal(size=(n, 7))
    X[:, :3] /= np.linalg.norm(X[:, :3], axis=1)[:, None]
    X[:, 3:] /= np.linalg.norm(X[:, 3:], axis=1)[:, None]
    T = np.stack([rng.uniform(0.25, 0.75, n), rng.uniform(0.3, 1.0, n)], 1)
    return X, T


def test_forward_and_gradient_match_autograd():
    rng = np.random.default_rng(0)
    p = rng.normal(size=mlp.NPARAM) * 0.3
    X, T = _batch(rng, 100)
    net = _torch_net(p)
    w = (1.0, 0.5)
    Y = net(torch.tensor(X))
    assert np.allclose(mlp.forward(p, X, bf=False), Y.detach().numpy(), rtol=1e-12, atol=1e-12)
    loss_t = ((Y - torch.tensor(T)) ** 2 * torch.tensor(w)).sum() / len(X)
    loss_t.backward()
    g_t = np.concatenate([np.concatenate([m.weight.grad.numpy().ravel(), m.bias.grad.numpy()])
                          for m in net if isinstance(m, torch.nn.Linear)])
    loss, g = mlp.loss_grad(p, X, T, w, bf=False)
    assert abs(loss - loss_t.item()) < 1e-12 * max(1.0, abs(loss))
    assert np.allclose(g, g_t, rtol=1e-10, atol=1e-12)


def test_gradient_central_differences():
    rng = np.random.default_rng(1)
    p = rng.normal(size=mlp.NPARAM) * 0.3
    X, T = _batch(rng, 20)
    _, g = mlp.loss_grad(p, X, T, bf=False)
    for q in rng.choice(mlp.NPARAM, 40, replace=False):
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[q] += h
        pm[q] -= h
        fd = (mlp.loss_grad(pp, X, T, bf=False)[0] - mlp.loss_grad(pm, X, T, bf=False)[0]) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(g[q])), q


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.normal(size=10000) * 10.0 ** rng.integers(-8, 8, 10000),
                        [0.0, -0.0, 1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -3.5]])
    ref = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mlp.bf16(x), ref)


def test_adam_step_matches_torch():
    rng = np.random.default_rng(3)
    p0 = rng.normal(size=50)
    grads = [rng.normal(size=50) for _ in range(5)]
    t = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([t], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for k, g in enumerate(grads, start=1):
        opt.zero_grad()
        t.grad = torch.tensor(g)
        opt.step()
        p, m, v = mlp.adam(p, m, v, g, k, 1e-2)
    assert np.allclose(p, t.detach().numpy(), rtol=1e-12, atol=1e-14)
