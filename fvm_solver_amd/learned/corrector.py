"""Learned window predictor with the plain Runge-Kutta solver as fallback and checker.

No reference counterpart.  Over a window of k solver steps of a (batched) 2-D channel
problem the corrector predicts the state at the window's midpoint and end from the
state U, the RHS f0 = f(U) and the RHS fm half a window earlier:

    U(t + W/2) = U + W (a_m f0 + b_m fm) + s * y_mid,
    U(t + W)   = U + W (a_e f0 + b_e fm) + s * y_end,

with W = k dt of the cell's mesh and y = MPNN(features).  The first terms integrate the
straight line through fm and f0 (an Adams-Bashforth extrapolation over the window, the
"base"); the network learns what the base misses, in units of s, per mesh the RMS over
its cells and conserved variables of W (f0 - fm).  The network sees, per cell, the normalised
primitives, W f0 and W (f0 - fm) in scaled units, two geometry features (log cell size,
share of boundary facets) and least-squares gradients of the primitives and of f0 times
the cell size: the direction in which the state is carried, which a neighbour mean alone
cannot see.

The history costs no extra RHS evaluation: every predicted window is checked with two
RHS evaluations, at the predicted midpoint and end,

    d = U1 - U - W/6 (f(U) + 4 f(Um) + f(U1)),     (Simpson)
    e = Um - (U + U1)/2 - W/8 (f(U) - f(U1)),       (Hermite midpoint)

both of which vanish to fourth order in W on a smooth exact trajectory, and when the
window is accepted f(Um) and f(U1) are the next window's fm and f0.  A window is
accepted for a mesh only if ||d|| / ||U1|| and ||e|| / ||U1|| over the mesh's whole
conserved state are at most tol.  tol bounds a window's defect, not the state error, which
grows over consecutive accepted windows: on 64 random channel meshes the default 3.5e-4
kept every mesh's state within 6.1e-4 of plain RK4 after 17 windows, where 1e-3 let the
largest meshes drift to 2e-3.  The meshes that fail are re-run with the plain fp32 RK4
solver (a masked step computes only the 256-cell tiles holding their cells), the others
keep their predictions, and f is evaluated at the merged midpoint and end to renew the
history.

After each window (accepted, partly re-run or plain) the solver takes ``settle`` plain RK4
steps before the next prediction (default 1): the scheme's own dissipation removes the
cell-scale roughness a prediction leaves, which the check amplifies and which would
otherwise build up over consecutive accepted windows.  The history fh is then f at the
previous window's midpoint, W/2 + settle dt before the new start, and the base integrates
the line through fh and f0 with that spacing (``base_coefficients``).

Training data are the batched solver's own trajectories: per window the network input
and the scaled residuals of the base at the true midpoint and end states.
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch

from fvm_solver_amd import _native as nat
from fvm_solver_amd.fvm_stepping.facet_process import FacetCalc
from fvm_solver_amd.learned.mpnn import MPNN

_BF = torch.bfloat16


def base_coefficients(k: int, settle: int):
    """(a, b) of the base U + W (a f0 + b fh) at a window's midpoint and end, for a
    history fh taken W/2 + settle dt before f0 (W = k dt): the integrals over [0, W/2]
    and [0, W] of the straight line through fh and f0, divided by W."""
    rho = 0.5 + settle / k                  # history spacing in units of W
    return (0.5 + 0.125 / rho, -0.125 / rho), (1.0 + 0.5 / rho, -0.5 / rho)



def lsq_neighbour_weights(dm) -> np.ndarray:
    """[NL][2][N] fp32 weights of the cells' interior face neighbours in the LSQ gradient
    rows (G_mats, slot order of ``cell_nbr``): grad v_c ~ sum_k w[k][:, c] (v_nb(k) - v_c),
    boundary faces left out (their weight is dropped)."""
    h = dm.host
    gval, nbr = h["gval"], h["nbr_s"]              # gval: [N, D, NL+1], slots then the cell
    N, NL = nbr.shape
    D = gval.shape[1]
    if D != 2:
        raise NotImplementedError("the learned corrector takes 2-D meshes")
    w = np.ascontiguousarray(gval[:, :, :NL].transpose(2, 1, 0)).astype(np.float32)
    w[np.broadcast_to((nbr >= N).T[:, None, :], w.shape)] = 0.0
    return w


class LearnedCorrector:
    """Corrector for one BatchedSolver (whose meshes, physics and dt it uses)."""

    KX = 32

    def __init__(self, solver, k: int = 8, hidden: int = 128, layers: int = 1, seed: int = 0, tol: float = 3.5e-4,
                 settle: int = 1):
        self.s = solver
        self.k = k
        self.tol = tol
        self.settle = settle
        self.base_mid, self.base_end = base_coefficients(k, settle)
        dm = solver.dm
        if dm.dim != 2:
            raise NotImplementedError("the learned corrector takes 2-D meshes")
        if k % 2:
            raise ValueError("the window must hold an even number of steps (its midpoint is predicted)")
        self.device = solver.device
        self.C = solver.C
        self.N = dm.N
        self.M = solver.n_meshes
        # a private RHS evaluator on the same device mesh (tile-fused when the tables
        # exist): the solver's own carried gradient state is not touched
        self.E = FacetCalc.from_device_mesh(dm, solver.E.cfg, self.device, fused=True)
        self.nbr = dm.t["cell_nbr"]                                   # int32 [NL, N]
        self.wg = torch.from_numpy(lsq_neighbour_weights(dm)).to(self.device)
        cm = dm.host["cell_mesh"].astype(np.int32)
        self.cell_mesh = torch.from_numpy(cm).to(self.device)
        self._cm_long = self.cell_mesh.long()
        self.off = torch.from_numpy(np.asarray(dm.cell_offsets, np.int64)).to(self.device)
        nb = dm.host["nbr_s"]
        bfrac = (nb >= self.N).sum(1) / nb.shape[1]
        logh = np.log(dm.h64.cpu().numpy())
        self.geo_raw = np.stack([logh, bfrac], 1)
        self._h = dm.h64.float().contiguous()
        self.net = MPNN(self.KX, 2 * self.C, hidden, layers, device=self.device, seed=seed)
        self.stats = None
        z = lambda *sh: torch.empty(*sh, device=self.device)
        N, C = self.N, self.C
        self._prim0 = z(N, C)          # primitives of the window start
        self._prim1 = z(N, C)
        self._f0 = z(N, C)             # f at the window start
        self._fh = z(N, C)             # f half a window earlier (history)
        self._fm = z(N, C)             # f at the predicted midpoint
        self._f1 = z(N, C)             # f at the predicted end
        self._Um = z(N, C)
        self._U1 = z(N, C)
        self._s1 = z(self.M, C)
        self._s2 = z(self.M, C)
        self._st = z(self.M, C)
        self._X = torch.empty(N, self.KX, dtype=_BF, device=self.device)
        self._def = torch.empty(self.M, 3 * C, dtype=torch.float64, device=self.device)
        self._part = torch.empty(self.M * 64 * 3 * C, dtype=torch.float64, device=self.device)
        self._hist_version = None      # solver version the history belongs to
        self.n_accept = 0
        self.n_reject = 0
        self.n_plain = 0               # plain windows that (re)started the history
        self.n_partial = 0             # windows with some meshes re-run
        self.mesh_accepted = 0         # accepted (mesh, window) predictions
        self.mesh_windows = 0          # checked (mesh, window) predictions
        self.last_defect = None

    # ------------------------------------------------------------------ pieces
    def window(self) -> torch.Tensor:
        return (self.k * self.s.dt[self._cm_long]).float().contiguous()

    def _state_to_prim(self, U, out):
        nat.check(nat.lib().fvm_state_to_prim_f32(nat.ptr(U), nat.ptr(out), self.N, self.C - 2, self.s.inv_cv,
                                                  nat.stream_ptr(self.device)), "fvm_state_to_prim_f32")

    def _rhs(self, U, prim, f):
        self._state_to_prim(U, prim)
        self.E.rhs(prim, out=f)

    def _rms(self, X, W, out):
        """Per mesh and variable, the RMS of W X over the mesh's cells."""
        nat.check(nat.lib().fvm_mesh_rms(nat.ptr(X), nat.ptr(W), nat.ptr(self.off), self.M, self.C, 1e-30,
                                         nat.ptr(out), nat.ptr(self._part), nat.stream_ptr(self.device)),
                  "fvm_mesh_rms")
        return out

    def _scales(self, f0, fh, W, s1, s2, st):
        """Per mesh and variable: s1 = RMS of W f0, s2 = RMS of W (f0 - fh); st = the
        target scale, per mesh the RMS of s2 over the variables (one scale for the whole
        state: the check and the error are relative L2 norms of the whole state, so the
        network's accuracy is spent where those norms are large)."""
        nat.check(nat.lib().fvm_mesh_rms2(nat.ptr(f0), nat.ptr(fh), nat.ptr(W), nat.ptr(self.off), self.M, self.C,
                                          1e-30, nat.ptr(s1), nat.ptr(s2), nat.ptr(self._part),
                                          nat.stream_ptr(self.device)), "fvm_mesh_rms2")
        st.copy_(torch.sqrt((s2 * s2).mean(1, keepdim=True)).expand(-1, self.C))

    def history_valid(self) -> bool:
        return self._hist_version == self.s.version

    def _settle_and_start(self):
        """The settle steps after a window, then f at the next window's start."""
        for _ in range(self.settle):
            self.s.step()
        self._rhs(self.s.state, self._prim0, self._f0)
        self._hist_version = self.s.version

    def plain_window(self, end_out: torch.Tensor | None = None):
        """k plain RK4 steps (their end state copied to end_out if given) and the settle
        steps, evaluating f at the window's midpoint (the history) and at the next
        window's start."""
        for _ in range(self.k // 2):
            self.s.step()
        self._rhs(self.s.state, self._prim1, self._fh)
        for _ in range(self.k // 2):
            self.s.step()
        if end_out is not None:
            end_out.copy_(self.s.state)
        self._settle_and_start()

    @staticmethod
    def _moments(chunks, floor_rel=1e-3, floor_abs=1e-8):
        """Per-column mean and (floored) standard deviation over a list of [n, C]
        tensors, accumulated in fp64."""
        n, s1, s2 = 0, 0.0, 0.0
        for x in chunks:
            xd = x.double()
            n += xd.shape[0]
            s1 = s1 + xd.sum(0)
            s2 = s2 + (xd * xd).sum(0)
        mu = s1 / n
        sd = torch.sqrt(torch.clamp(s2 / n - mu * mu, min=0.0))
        sd = torch.maximum(sd, floor_rel * (mu.abs() + sd))
        sd = torch.clamp(sd, min=floor_abs)
        return mu.float().cpu(), sd.float().cpu()

    def set_stats_from(self, prims):
        """Normalisation of the primitive inputs and the geometry features."""
        mu_p, sd_p = self._moments(prims)
        g = torch.from_numpy(self.geo_raw).float()
        self.stats = {"mu_p": mu_p, "sd_p": sd_p, "mu_g": g.mean(0), "sd_g": g.std(0).clamp(min=1e-6)}
        self._apply_stats()

    def set_stats(self, U):
        """Normalisation from a single representative state."""
        self._state_to_prim(U, self._prim1)
        self.set_stats_from([self._prim1])

    def _apply_stats(self):
        st = self.stats
        self._mu_p = st["mu_p"].to(self.device)
        self._sd_p = st["sd_p"].to(self.device)
        g = (torch.from_numpy(self.geo_raw).float() - st["mu_g"]) / st["sd_g"]
        self._geo = g.contiguous().to(self.device)

    def _features_from(self, prim, f0, fh, W, s1, s2, X):
        nat.check(nat.lib().fvm_mp_window_features(nat.ptr(prim), nat.ptr(f0), nat.ptr(fh), nat.ptr(W),
                                                   nat.ptr(self._geo), nat.ptr(self._h), nat.ptr(self.nbr),
                                                   self.nbr.shape[0], nat.ptr(self.wg), self.N, self.C,
                                                   nat.ptr(self._mu_p), nat.ptr(self._sd_p), nat.ptr(s1), nat.ptr(s2),
                                                   nat.ptr(self.cell_mesh), nat.ptr(X), self.KX,
                                                   nat.stream_ptr(self.device)), "fvm_mp_window_features")
        return X

    def _apply(self, U, f0, fh, W, coef, y, col, scale, out):
        a, b = coef
        nat.check(nat.lib().fvm_mp_window_apply(nat.ptr(U), nat.ptr(f0), nat.ptr(fh), a, b, nat.ptr(W),
                                                nat.ptr(None if y is None else y[:, col:]), 2 * self.C,
                                                nat.ptr(scale), nat.ptr(self.cell_mesh), self.N, self.C, nat.ptr(out),
                                                nat.stream_ptr(self.device)), "fvm_mp_window_apply")
        return out

    # ------------------------------------------------------------------ data
    def generate(self, n_windows: int, log=None):
        """Run the solver for n_windows windows (after one plain window that starts the
        history) and keep, per window, the network input and the scaled residuals of
        the base at the true midpoint and end; the solver's state advances.  Returns the
        dataset list of (X, target [N, 2C])."""
        W = self.window()
        if not self.history_valid():
            self.plain_window()
        raw = []
        for i in range(n_windows):
            U0 = self.s.state.clone()
            prim, f0, fh = self._prim0.clone(), self._f0.clone(), self._fh.clone()
            s1, s2, st = torch.empty_like(self._s1), torch.empty_like(self._s2), torch.empty_like(self._st)
            self._scales(f0, fh, W, s1, s2, st)
            for _ in range(self.k // 2):
                self.s.step()
            rm = self.s.state - self._apply(U0, f0, fh, W, self.base_mid, None, 0, None, self._Um)
            self._rhs(self.s.state, self._prim1, self._fh)
            for _ in range(self.k // 2):
                self.s.step()
            r1 = self.s.state - self._apply(U0, f0, fh, W, self.base_end, None, 0, None, self._U1)
            self._settle_and_start()
            raw.append((prim, f0, fh, s1, s2, st, torch.cat([rm, r1], 1)))
            del U0
            if log and (i + 1) % max(1, n_windows // 5) == 0:
                log(f"generated {i + 1}/{n_windows} windows")
        self.set_stats_from([r[0] for r in raw])
        data = []
        while raw:
            prim, f0, fh, s1, s2, st, r = raw.pop(0)
            X = torch.empty(self.N, self.KX, dtype=_BF, device=self.device)
            sc = st[self._cm_long].repeat(1, 2)
            data.append((self._features_from(prim, f0, fh, W, s1, s2, X), (r / sc).contiguous()))
        return data

    # ------------------------------------------------------------------ training
    def train(self, data, iters: int, lr: float = 2e-3, lr_min: float = 1e-4, seed: int = 0, log=None):
        """Adam on random windows with a cosine learning-rate schedule; returns the loss
        history (device values are fetched every 50 iterations)."""
        rng = np.random.default_rng(seed)
        hist = []
        for it in range(iters):
            X, T = data[int(rng.integers(len(data)))]
            cur = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * it / max(iters - 1, 1)))
            loss = self.net.train_step(X, self.nbr, T, lr=cur)
            if it % 50 == 0 or it == iters - 1:
                hist.append((it, float(loss)))
                if log:
                    log(f"iter {it}: loss {hist[-1][1]:.4e}")
        return hist

    def eval_loss(self, data) -> float:
        tot = 0.0
        for X, T in data:
            self.net.forward(X, self.nbr, train=True)
            tot += float(self.net.loss(T))
        return tot / len(data)

    # ------------------------------------------------------------------ inference
    def predict(self, use_net: bool = True):
        """Predicted midpoint and end states of the window from the solver's state (into
        the corrector's buffers), with the window lengths.  use_net=False gives the base
        alone.  Needs a valid history."""
        if not self.history_valid():
            raise RuntimeError("the history does not belong to the solver's state: run plain_window() first")
        U0 = self.s.state
        W = self.window()
        self._scales(self._f0, self._fh, W, self._s1, self._s2, self._st)
        y = None
        if use_net:
            X = self._features_from(self._prim0, self._f0, self._fh, W, self._s1, self._s2, self._X)
            y = self.net.forward(X, self.nbr)
        (am, bm), (ae, be) = self.base_mid, self.base_end
        nat.check(nat.lib().fvm_mp_window_predict(nat.ptr(U0), nat.ptr(self._f0), nat.ptr(self._fh), am, bm, ae, be,
                                                  nat.ptr(W), nat.ptr(y), 2 * self.C, nat.ptr(self._st),
                                                  nat.ptr(self.cell_mesh), self.N, self.C, nat.ptr(self._Um),
                                                  nat.ptr(self._U1), nat.stream_ptr(self.device)),
                  "fvm_mp_window_predict")
        return self._Um, self._U1, W

    def defect(self, U0, Um, U1, W) -> np.ndarray:
        """Per mesh, max(||d||, ||e||) / ||U1|| of the Simpson and Hermite-midpoint
        defects over the mesh's whole state (host array), with f(U0) = the history's f0;
        f(Um) and f(U1) are left in the corrector's buffers.  Per-variable ratios are
        kept in ``last_defect``."""
        self._rhs(Um, self._prim1, self._fm)
        self._rhs(U1, self._prim1, self._f1)
        nat.check(nat.lib().fvm_mp_defect3(nat.ptr(U0), nat.ptr(Um), nat.ptr(U1), nat.ptr(self._f0),
                                           nat.ptr(self._fm), nat.ptr(self._f1), nat.ptr(W), nat.ptr(self.off),
                                           self.M, self.C, nat.ptr(self._def), nat.ptr(self._part),
                                           nat.stream_ptr(self.device)), "fvm_mp_defect3")
        d = self._def.cpu().numpy()
        C = self.C
        u2 = np.maximum(d[:, 2 * C:].sum(1), 1e-300)
        self.last_defect = np.sqrt(d[:, :C] / np.maximum(d[:, 2 * C:], 1e-300))
        return np.sqrt(np.maximum(d[:, :C].sum(1), d[:, C:2 * C].sum(1)) / u2)

    def advance(self, use_net: bool | str = True) -> bool:
        """One window.  use_net True: predict with the network and check; "base": the
        same with the base predictor alone; False: plain RK4.  Meshes whose prediction
        fails the check are re-run with plain RK4 (only the tiles holding their cells are
        stepped, the other meshes keep their accepted predictions); a window without a
        valid history is run plainly.  Returns True when every mesh's prediction was
        accepted."""
        if use_net is False:
            for _ in range(self.k + self.settle):
                self.s.step()
            return False
        if not self.history_valid():
            self.n_plain += 1
            self.plain_window()
            return False
        U0 = self.s.state
        Um, U1, W = self.predict(use_net is True)
        ok = self.defect(U0, Um, U1, W) <= self.tol
        self.mesh_accepted += int(ok.sum())
        self.mesh_windows += self.M
        if ok.all():
            self.s.set_state(U1)
            # f at the accepted midpoint is the next window's history; without settle
            # steps f at the accepted end is the next start's f
            self._fh, self._fm = self._fm, self._fh
            if self.settle == 0:
                self._f0, self._f1 = self._f1, self._f0
                self._prim0, self._prim1 = self._prim1, self._prim0
                self._hist_version = self.s.version
            else:
                self._settle_and_start()
            self.n_accept += 1
            return True
        if not ok.any() or not self.s.E.fused:
            self.n_reject += 1
            self.plain_window()
            return False
        self.n_partial += 1
        self._partial_window(ok)
        return False

    def _partial_window(self, ok: np.ndarray):
        """Plain RK4 for the meshes with ok false, predictions for the others; the
        history is evaluated on the merged midpoint and end states."""
        tiles = self.s.tile_mask(~ok)
        redo = torch.from_numpy(~ok).to(self.device)[self.s.cell_mesh][:, None]
        for _ in range(self.k // 2):
            self.s.step(tiles)
        mid = torch.where(redo, self.s.state, self._Um)
        self._rhs(mid, self._prim1, self._fh)
        for _ in range(self.k // 2):
            self.s.step(tiles)
        self.s.set_state(torch.where(redo, self.s.state, self._U1))
        self._settle_and_start()

    def run(self, n_windows: int, use_net: bool | str = True):
        """n_windows windows; returns (seconds, accepted windows)."""
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        acc = sum(self.advance(use_net) for _ in range(n_windows))
        torch.cuda.synchronize()
        return time.perf_counter() - t0, acc
