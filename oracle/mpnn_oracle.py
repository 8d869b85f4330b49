"""CPU oracle of the learned corrector (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline may import this file;
the product path never does.

The corrector has no counterpart in the reference (it defines no neural network), so
this restatement follows the build's own specification, written independently of the
HIP kernels:

    h0      = relu(x W_in + b_in)
    a_l     = [h_l | mean over the cell's interior face neighbours of h_l]
    h_{l+1} = h_l + relu(a_l W_l + b_l)                        l = 0 .. L-1
    y       = h_L W_out + b_out
    loss    = mean((y - t)^2)

in float64 torch on the CPU, with gradients from torch.autograd and Adam as in
torch.optim.Adam.  Optionally the activations are rounded to bfloat16 at the points the
GPU path stores them (``bf16_points=True``), so that the two agree to fp32 accumulation
order rather than to bf16 resolution.

What pins it (tests/test_mpnn_oracle.py): the aggregation equals a dense
row-normalised adjacency product built independently from the cell lists; the
autograd gradients agree with central finite differences (torch.autograd.gradcheck);
the Adam step equals torch.optim.Adam's.
"""
from __future__ import annotations

import numpy as np
import torch

def param_names(layers: int = 3):
    return ["W_in", "b_in"] + [f"{k}_{l}" for l in range(layers) for k in ("W", "b")] + ["W_out", "b_out"]


def _bf(x: torch.Tensor) -> torch.Tensor:
    return x.to(torch.bfloat16).to(x.dtype)


def neighbour_mean(h: torch.Tensor, nbr: np.ndarray) -> torch.Tensor:
    """mean over interior neighbours; nbr [NL][N] with entries >= N marking boundary
    facets (no neighbour).  Cells without a neighbour get 0."""
    N = h.shape[0]
    nbr = torch.as_tensor(np.asarray(nbr, np.int64))
    valid = nbr < N
    cnt = valid.sum(0).to(h.dtype)
    s = torch.zeros_like(h)
    for k in range(nbr.shape[0]):
        idx = torch.where(valid[k], nbr[k], torch.zeros_like(nbr[k]))
        s = s + h[idx] * valid[k].to(h.dtype)[:, None]
    return s / torch.clamp(cnt, min=1.0)[:, None]


def dense_mean_operator(nbr: np.ndarray) -> np.ndarray:
    """Row-normalised adjacency matrix (dense, fp64) from the cell lists: an
    independent construction used to pin ``neighbour_mean``."""
    NL, N = nbr.shape
    A = np.zeros((N, N))
    for c in range(N):
        ns = [int(nbr[k, c]) for k in range(NL) if nbr[k, c] < N]
        for n in ns:
            A[c, n] += 1.0 / len(ns)
    return A


def forward(params: dict, x: torch.Tensor, nbr: np.ndarray, layers: int = 3, bf16_points: bool = False):
    """y = MPNN(x).  params: name -> float64 tensors (W as [in, out])."""
    r = _bf if bf16_points else (lambda t: t)
    h = r(torch.relu(x @ params["W_in"] + params["b_in"]))
    for l in range(layers):
        a = torch.cat([h, r(neighbour_mean(h, nbr))], 1)
        h = r(h + torch.relu(a @ params[f"W_{l}"] + params[f"b_{l}"]))
    return h @ params["W_out"] + params["b_out"]


def loss_and_grads(params: dict, x: torch.Tensor, t: torch.Tensor, nbr: np.ndarray, layers: int = 3,
                   bf16_points: bool = False):
    p = {k: v.detach().clone().requires_grad_(True) for k, v in params.items()}
    y = forward(p, x, nbr, layers, bf16_points)
    loss = ((y - t) ** 2).mean()
    loss.backward()
    return float(loss), {k: v.grad.detach() for k, v in p.items()}, y.detach()


def adam_step(w, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (returns new w, m, v); torch.optim.Adam's arithmetic."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w - lr * mh / (torch.sqrt(vh) + eps), m, v


def init_params(kx: int, hidden: int, n_out: int, layers: int = 3, seed: int = 0, dtype=torch.float64):
    """Uniform(+-sqrt(6 / fan_in)) weights, zero biases, small output layer."""
    g = torch.Generator().manual_seed(seed)

    def u(fan_in, shape, gain=1.0):
        a = gain * (6.0 / fan_in) ** 0.5
        return (torch.rand(shape, generator=g, dtype=dtype) * 2 - 1) * a

    p = {"W_in": u(kx, (kx, hidden)), "b_in": torch.zeros(hidden, dtype=dtype)}
    for l in range(layers):
        p[f"W_{l}"] = u(2 * hidden, (2 * hidden, hidden), 0.5)
        p[f"b_{l}"] = torch.zeros(hidden, dtype=dtype)
    p["W_out"] = u(hidden, (hidden, n_out), 0.1)
    p["b_out"] = torch.zeros(n_out, dtype=dtype)
    return p
