"""Pins of the corrector's CPU oracle (no GPU): aggregation against an independent dense
construction, autograd gradients against finite differences, Adam against torch.optim."""
import numpy as np
import torch

from oracle import mpnn_oracle as MO


def _graph(n=40, seed=0):
    """Random symmetric graph with <= 3 neighbours per node, padded with N + b entries."""
    rng = np.random.default_rng(seed)
    nbr = np.full((3, n), -1, np.int64)
    deg = np.zeros(n, np.int64)
    for _ in range(4 * n):
        a, b = rng.integers(0, n, 2)
        if a == b or deg[a] >= 3 or deg[b] >= 3 or b in nbr[:, a]:
            continue
        nbr[deg[a], a] = b
        nbr[deg[b], b] = a
        deg[a] += 1
        deg[b] += 1
    pad = nbr < 0
    nbr[pad] = n + np.arange(pad.sum())
    return nbr


def test_neighbour_mean_matches_dense_operator():
    nbr = _graph()
    h = torch.randn(nbr.shape[1], 8, dtype=torch.float64)
    A = torch.from_numpy(MO.dense_mean_operator(nbr))
    assert torch.allclose(MO.neighbour_mean(h, nbr), A @ h, atol=1e-14)
    # the operator's rows sum to one where a node has neighbours, zero otherwise
    rs = A.sum(1).numpy()
    has = (nbr < nbr.shape[1]).any(0)
    assert np.allclose(rs[has], 1.0) and np.allclose(rs[~has], 0.0)


def test_gradients_match_finite_differences():
    nbr = _graph(12, 1)
    kx, H, C = 5, 6, 3
    p = MO.init_params(kx, H, C, seed=3)
    x = torch.randn(12, kx, dtype=torch.float64)
    t = torch.randn(12, C, dtype=torch.float64)
    names = MO.param_names()
    _, g, _ = MO.loss_and_grads(p, x, t, nbr)

    def f(*ws):
        q = dict(zip(names, ws))
        return ((MO.forward(q, x, nbr) - t) ** 2).mean()

    ws = tuple(p[k].clone().requires_grad_(True) for k in names)
    assert torch.autograd.gradcheck(f, ws, eps=1e-6, atol=1e-7)
    # and a planted mistake in the aggregation changes the output
    y_ok = MO.forward(p, x, nbr)
    bad = nbr.copy()
    bad[0] = bad[1]
    assert not torch.allclose(MO.forward(p, x, bad), y_ok)
    assert set(g) == set(names)


def test_adam_step_matches_torch_optim():
    w0 = torch.randn(50, dtype=torch.float64)
    w = w0.clone()
    m = torch.zeros_like(w)
    v = torch.zeros_like(w)
    wt = torch.nn.Parameter(w0.clone())
    opt = torch.optim.Adam([wt], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    for step in range(1, 6):
        g = torch.sin(w * step)
        w, m, v = MO.adam_step(w, g, m, v, step, lr=3e-3)
        opt.zero_grad()
        wt.grad = torch.sin(wt.detach() * step)
        opt.step()
        assert torch.allclose(w, wt.detach(), rtol=1e-12, atol=1e-14)
