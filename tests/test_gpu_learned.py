"""The learned window corrector on GPU: trained on the batched solver's own trajectories,
it must beat the method without it (the base extrapolation alone), its accepted windows
must pass the defect check, and with a corrupted network every window must fall back
to the plain RK4 solver, reproducing the plain run exactly."""
import numpy as np
import pytest
import torch

from fvm_solver_amd.batch import BatchedSolver, random_channel_batch
from fvm_solver_amd.learned.corrector import LearnedCorrector, lsq_neighbour_weights

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


def test_lsq_weights_are_exact_on_linear_fields():
    """The feature gradients: interior cells reproduce the gradient of a linear field."""
    s, problems, _ = _solver(5, n=2)
    w = lsq_neighbour_weights(s.dm)
    nbr = s.dm.host["nbr_s"]
    N = nbr.shape[0]
    cen = np.concatenate([m.geo64["centroids"] for m, _, _, _ in problems])
    v = 0.3 * cen[:, 0] - 1.7 * cen[:, 1]
    interior = (nbr < N).all(1)
    g = np.zeros((N, 2))
    for k in range(nbr.shape[1]):
        nb = np.where(nbr[:, k] < N, nbr[:, k], 0)
        g += w[k].T * (v[nb] - v)[:, None] * (nbr[:, k] < N)[:, None]
    assert np.allclose(g[interior], [0.3, -1.7], atol=1e-4)


def test_training_beats_base_window(trained):
    lc, data, loss0, loss1 = trained
    # predicting zero correction (the base extrapolation alone) scores mean(T^2) in the
    # scaled units; the trained network must explain a good part of it on windows it
    # never saw
    zero = float(np.mean([float((T.double() ** 2).mean()) for _, T in data[-8:]]))
    print(f"[gpu-test]   base-window loss {zero:.4f}, trained {loss1:.4f}")
    assert loss1 < 0.8 * zero


def test_prediction_error_vs_rk4(trained):
    """Continue the training trajectory: compare one predicted window with k RK4 steps."""
    lc, _, _, _ = trained
    s = lc.s
    assert lc.history_valid()
    U0 = s.state.clone()
    _, U1, W = lc.predict(use_net=True)
    U1 = U1.clone()
    _, Ub, _ = lc.predict(use_net=False)
    Ub = Ub.clone()
    a, b = lc.base_end
    assert torch.allclose(Ub, U0 + W[:, None] * (a * lc._f0 + b * lc._fh), rtol=1e-6, atol=1e-6)
    for _ in range(lc.k):
        s.step()
    ref = s.state
    inc = (ref - U0).double()
    e_net = float((U1 - ref).double().norm() / inc.norm())
    e_base = float((Ub - ref).double().norm() / inc.norm())
    per_var = ((U1 - ref).double().norm(dim=0) / inc.norm(dim=0)).cpu().numpy()
    print(f"[gpu-test]   window increment error: net {e_net:.4f} (per variable {per_var}), base {e_base:.4f}")
    assert e_net < e_base


def test_accepted_windows_pass_the_check(trained):
    """Windows the corrector accepts keep the state close to the plain solver's."""
    lc, _, _, _ = trained
    s = lc.s
    lc.plain_window()
    U0 = s.state.clone()
    acc = sum(lc.advance() for _ in range(4))
    Ul = s.state.clone()
    s.set_state(U0)
    for _ in range(4 * (lc.k + lc.settle)):
        s.step()
    err = float((Ul - s.state).double().norm() / s.state.double().norm())
    print(f"[gpu-test]   accepted {acc}/4 windows, state rel-L2 vs plain after 4 windows {err:.2e}")
    assert err < 4 * lc.tol


def test_corrupted_network_falls_back_to_plain_rk4():
    s, problems, _ = _solver(21, n=3)
    ref = BatchedSolver(problems, device=DEV)
    ref.set_dt(s.dt.cpu().numpy())
    lc = LearnedCorrector(s, k=4, seed=1)
    lc.set_stats(s.state)
    lc.net.w["b_out"].fill_(1e3)               # a prediction far off the solution
    for _ in range(3):
        assert lc.advance() is False
        for _ in range(lc.k + lc.settle):
            ref.step()
    torch.cuda.synchronize()
    assert lc.n_plain == 1 and lc.n_reject == 2 and lc.n_accept == 0
    assert torch.equal(s.state, ref.state)


def test_masked_step_equals_full_step_on_selected_meshes():
    s, problems, _ = _solver(41, n=5)
    ref = BatchedSolver(problems, device=DEV)
    ref.set_dt(s.dt.cpu().numpy())
    sel = np.array([True, False, True, False, True])
    tiles = s.tile_mask(sel)
    assert tiles.dtype == torch.int32 and bool((tiles[1:] > tiles[:-1]).all())
    with pytest.raises(ValueError):
        s.tile_mask(np.zeros(5, bool))          # an empty list would step every mesh
    for _ in range(6):
        s.step(tiles)
        ref.step()
    torch.cuda.synchronize()
    for i in np.flatnonzero(sel):
        assert torch.equal(s.state_of(i), ref.state_of(i)), i


def test_partial_window_merges_predictions_and_plain_steps():
    """Meshes that fail the check take the plain RK4 window (bit for bit), the others
    keep their predicted end state."""
    s, problems, _ = _solver(31, n=4)
    ref = BatchedSolver(problems, device=DEV)
    ref.set_dt(s.dt.cpu().numpy())
    lc = LearnedCorrector(s, k=4, seed=2, settle=0)
    lc.set_stats(s.state)
    lc.plain_window()
    for _ in range(lc.k):
        ref.step()
    assert torch.equal(s.state, ref.state)
    Um, U1, W = lc.predict(use_net=False)
    rel = lc.defect(s.state, Um, U1, W)
    ok = rel <= np.median(rel)
    assert ok.any() and not ok.all()
    U1p = U1.clone()
    lc._partial_window(ok)
    for _ in range(lc.k):
        ref.step()
    torch.cuda.synchronize()
    assert lc.history_valid()
    for i in range(4):
        lo, hi = int(s.offsets[i]), int(s.offsets[i + 1])
        want = U1p[lo:hi] if ok[i] else ref.state_of(i)
        assert torch.equal(s.state_of(i), want), i



def test_window_kernels_against_torch():
    """The window's scale, feature and predict kernels against torch restatements of
    their formulas (fp32; the features to bf16 storage precision)."""
    from fvm_solver_amd import _native as nat
    s, problems, _ = _solver(17, n=3)
    lc = LearnedCorrector(s, k=8, seed=4)
    lc.set_stats(s.state)
    lc.plain_window()
    W = lc.window()
    M, C, N = lc.M, lc.C, lc.N
    lc._scales(lc._f0, lc._fh, W, lc._s1, lc._s2, lc._st)
    cm = lc._cm_long

    def mesh_rms(x):
        ss = torch.zeros(M, C, dtype=torch.float64, device=DEV).index_add_(0, cm, x.double() ** 2)
        n = torch.bincount(cm, minlength=M).double()[:, None]
        return torch.sqrt(ss / n)

    f0, fh, prim = lc._f0, lc._fh, lc._prim0
    s1 = mesh_rms(W[:, None] * f0)
    s2 = mesh_rms(W[:, None] * (f0 - fh))
    assert torch.allclose(lc._s1.double(), s1, rtol=1e-5)
    assert torch.allclose(lc._s2.double(), s2, rtol=1e-5)
    assert torch.allclose(lc._st.double(), torch.sqrt((s2 ** 2).mean(1, keepdim=True)).expand(-1, C), rtol=1e-5)

    X = lc._features_from(prim, f0, fh, W, lc._s1, lc._s2, lc._X).float()
    nbr = lc.nbr.long()
    h = lc._h[:, None]
    grad = lambda v: torch.stack([sum(torch.where((nbr[k] < N)[:, None], lc.wg[k, i][:, None] *
                                                   (v[nbr[k].clamp(max=N - 1)] - v), torch.zeros_like(v))
                                      for k in range(nbr.shape[0])) for i in range(2)])
    gp, gf = grad(prim), grad(f0)
    s1c, s2c = lc._s1[cm], lc._s2[cm]
    ref = torch.cat([(prim - lc._mu_p) / lc._sd_p, W[:, None] * f0 / s1c, W[:, None] * (f0 - fh) / s2c, lc._geo,
                     h * gp[0] / lc._sd_p, h * gp[1] / lc._sd_p, h * W[:, None] * gf[0] / s1c,
                     h * W[:, None] * gf[1] / s1c, torch.zeros(N, 2, device=DEV)], 1)
    err = (X - ref).abs() / (ref.abs() + 1e-2)
    print(f"[gpu-test]   window features: max rel err {float(err.max()):.2e}")
    assert float(err.max()) < 2e-2

    y = torch.randn(N, 2 * C, device=DEV, generator=torch.Generator(device=DEV).manual_seed(3))
    Um, U1 = torch.empty_like(f0), torch.empty_like(f0)
    (am, bm), (ae, be) = lc.base_mid, lc.base_end
    U = s.state
    nat.check(nat.lib().fvm_mp_window_predict(nat.ptr(U), nat.ptr(f0), nat.ptr(fh), am, bm, ae, be, nat.ptr(W),
                                              nat.ptr(y), 2 * C, nat.ptr(lc._st), nat.ptr(lc.cell_mesh), N, C,
                                              nat.ptr(Um), nat.ptr(U1), nat.stream_ptr(DEV)), "predict")
    st = lc._st[cm]
    assert torch.allclose(Um, U + W[:, None] * (am * f0 + bm * fh) + st * y[:, :C], rtol=1e-5, atol=1e-5)
    assert torch.allclose(U1, U + W[:, None] * (ae * f0 + be * fh) + st * y[:, C:], rtol=1e-5, atol=1e-5)
