"""The learned window corrector on GPU: trained on the batched solver's own trajectories,
it must beat the method without it (the bare Euler window), its accepted windows must
pass the defect check, and with a corrupted network every window must fall back to
the plain RK4 solver, reproducing the plain run exactly."""
import numpy as np
import pytest
import torch

from fvm_solver_amd.batch import BatchedSolver, random_channel_batch
from fvm_solver_amd.learned.corrector import LearnedCorrector

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _solver(seed, n=4, n_min=3000, n_max=8000):
    problems, params = random_channel_batch(n, seed=seed, n_min=n_min, n_max=n_max, device=DEV)
    s = BatchedSolver(problems, device=DEV)
    s.set_dt(s.cfl_dt(0.5))
    return s, problems, params


@pytest.fixture(scope="module")
def trained():
    s, problems, _ = _solver(11)
    lc = LearnedCorrector(s, k=8, seed=0)
    data = lc.generate(48)
    loss0 = lc.eval_loss(data[-8:])
    hist = lc.train(data[:-8], iters=600, lr=2e-3)
    loss1 = lc.eval_loss(data[-8:])
    print(f"[gpu-test]   held-out windows: loss untrained {loss0:.4f} -> trained {loss1:.4f}; "
          f"train loss {hist[0][1]:.4f} -> {hist[-1][1]:.4f}")
    return lc, data, loss0, loss1


def test_training_beats_euler_window(trained):
    lc, data, loss0, loss1 = trained
    # predicting zero correction (the bare Euler steps) scores mean(T^2) in the scaled
    # units; the trained network must explain a good part of it on windows it never saw
    zero = float(np.mean([float((T.double() ** 2).mean()) for _, T in data[-8:]]))
    print(f"[gpu-test]   Euler-window loss {zero:.4f}, trained {loss1:.4f}")
    assert loss1 < 0.7 * zero


def test_prediction_error_vs_rk4(trained):
    """Continue the training trajectory: compare one predicted window with k RK4 steps."""
    lc, _, _, _ = trained
    s = lc.s
    U0 = s.state.clone()
    Um, U1, f0, W = lc.predict(U0)
    U1 = U1.clone()
    euler = U0 + W[:, None] * f0
    for _ in range(lc.k):
        s.step()
    ref = s.state
    inc = (ref - U0).double()
    e_net = float((U1 - ref).double().norm() / inc.norm())
    e_eul = float((euler - ref).double().norm() / inc.norm())
    per_var = ((U1 - ref).double().norm(dim=0) / inc.norm(dim=0)).cpu().numpy()
    print(f"[gpu-test]   window increment error: net {e_net:.4f} (per variable {per_var}), Euler {e_eul:.4f}")
    assert e_net < e_eul


def test_corrupted_network_falls_back_to_plain_rk4():
    s, problems, _ = _solver(21, n=3)
    ref = BatchedSolver(problems, device=DEV)
    ref.set_dt(s.dt.cpu().numpy())
    lc = LearnedCorrector(s, k=4, seed=1)
    lc.set_stats(s.state)
    lc.net.w["b_out"].fill_(1e3)               # a prediction far off the solution
    for _ in range(3):
        assert lc.advance() is False
        for _ in range(lc.k):
            ref.step()
    torch.cuda.synchronize()
    assert lc.n_reject == 3 and lc.n_accept == 0
    assert torch.equal(s.state, ref.state)
