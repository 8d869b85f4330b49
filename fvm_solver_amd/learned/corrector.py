"""Learned window predictor with the plain Runge-Kutta solver as fallback and judge.

No reference counterpart.  Over a window of k solver steps of a (batched) channel
problem the corrector predicts the state at the window's midpoint and end,

    U(t + W/2) = U + (W/2) f(U) + s * y_mid,     U(t + W) = U + W f(U) + s * y_end,

with W = k dt of the cell's mesh and y = MPNN(features): explicit Euler steps plus a
learned correction, where s is, per mesh and conserved variable, the RMS of the Euler
increment W f(U) over the mesh's cells.  The network therefore learns the window's
residual relative to the size of the window's change, which is what varies least
between an impulsive start and a developed flow.  Inputs per cell: the normalised
primitives, the Euler increment divided by s, two geometry features (log cell size,
share of boundary facets) and the limited gradient of the RHS evaluation times the cell
size; the graph is the mesh's face adjacency.

Every predicted window is checked before it is accepted, with two more RHS evaluations
(at the predicted midpoint and end): the Simpson defect

    d = U1 - U - W/6 (f(U) + 4 f(Um) + f(U1))

and the Hermite-midpoint defect e = Um - (U + U1)/2 - W/8 (f(U) - f(U1)).  Both vanish to
fourth order in W for a smooth exact trajectory.  A window is accepted only if, on every
mesh of the batch, ||d|| / ||U1|| and ||e|| / ||U1|| over the mesh's whole conserved
state are at most tol.  Otherwise it is discarded and re-run with the plain fp32 RK4
solver.

Training data are the batched solver's own trajectories: (U_t, U_{t+k/2}, U_{t+k})
triples of every mesh of the batch, stored as the network's inputs and the scaled
residuals of the two Euler steps.
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


class LearnedCorrector:
    """Corrector for one BatchedSolver (whose meshes, physics and dt it uses)."""

    KX = 32

    def __init__(self, solver, k: int = 8, hidden: int = 128, layers: int = 3, seed: int = 0, tol: float = 1e-3):
        self.s = solver
        self.k = k
        self.tol = tol
        dm = solver.dm
        self.device = solver.device
        self.C = solver.C
        self.N = dm.N
        self.M = solver.n_meshes
        # a private RHS evaluator on the same device mesh: the solver's own carried
        # gradient state is not touched by the corrector's evaluations
        self.E = FacetCalc.from_device_mesh(dm, solver.E.cfg, self.device, fused=False)
        self.nbr = dm.t["cell_nbr"]                                   # int32 [NL, N]
        cm = dm.host["cell_mesh"].astype(np.int32)
        self.cell_mesh = torch.from_numpy(cm).to(self.device)
        self._cm_long = self.cell_mesh.long()
        self.off = torch.from_numpy(np.asarray(dm.cell_offsets, np.int64)).to(self.device)
        nb = dm.host["nbr_s"]
        bfrac = (nb >= self.N).sum(1) / nb.shape[1]
        logh = np.log(dm.h64.cpu().numpy())
        self.geo_raw = np.stack([logh, bfrac], 1)
        self._h = dm.h64.float().contiguous()
        if k % 2:
            raise ValueError("the window must hold an even number of steps (its midpoint is predicted)")
        self.net = MPNN(self.KX, 2 * self.C, hidden, layers, device=self.device, seed=seed)
        self.stats = None
        z = lambda *sh: torch.empty(*sh, device=self.device)
        self._prim = z(self.N, self.C)
        self._prim1 = z(self.N, self.C)
        self._f0 = z(self.N, self.C)
        self._fm = z(self.N, self.C)
        self._f1 = z(self.N, self.C)
        self._Um = z(self.N, self.C)
        self._U1 = z(self.N, self.C)
        self._sc = z(self.M, self.C)
        self._X = torch.empty(self.N, self.KX, dtype=_BF, device=self.device)
        self._def = torch.empty(self.M, 3 * self.C, dtype=torch.float64, device=self.device)
        self._part = torch.empty(self.M * 64 * 3 * self.C, dtype=torch.float64, device=self.device)
        self.n_accept = 0
        self.n_reject = 0
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

    def _inc_scale(self, f0, W, out):
        """Per mesh and variable, the RMS of the Euler increment W f over the mesh."""
        nat.check(nat.lib().fvm_mesh_rms(nat.ptr(f0), nat.ptr(W), nat.ptr(self.off), self.M, self.C, 1e-30,
                                         nat.ptr(out), nat.ptr(self._part), nat.stream_ptr(self.device)),
                  "fvm_mesh_rms")
        return out

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
        self._rhs(U, self._prim, self._f0)
        self.set_stats_from([self._prim])

    def _apply_stats(self):
        st = self.stats
        self._mu_p = st["mu_p"].to(self.device)
        self._sd_p = st["sd_p"].to(self.device)
        g = (torch.from_numpy(self.geo_raw).float() - st["mu_g"]) / st["sd_g"]
        self._geo = g.contiguous().to(self.device)

    def _features_from(self, prim, f0, grad, W, sc, X):
        nat.check(nat.lib().fvm_mp_features(nat.ptr(prim), nat.ptr(f0), nat.ptr(W), nat.ptr(self._geo), 2,
                                            nat.ptr(grad), nat.ptr(self._h), self.C - 2, self.N, self.C,
                                            nat.ptr(self._mu_p), nat.ptr(self._sd_p), nat.ptr(sc),
                                            nat.ptr(self.cell_mesh), nat.ptr(X), self.KX,
                                            nat.stream_ptr(self.device)), "fvm_mp_features")
        return X

    def features(self, U):
        """prim, f(U), the increment scale and the network input of state U (into the
        corrector's buffers)."""
        self._rhs(U, self._prim, self._f0)
        W = self.window()
        self._inc_scale(self._f0, W, self._sc)
        self._features_from(self._prim, self._f0, self.E.gp, W, self._sc, self._X)
        return self._X, self._f0, W, self._sc

    # ------------------------------------------------------------------ data
    def generate(self, n_windows: int, log=None):
        """Run the solver for n_windows windows and keep, per window, the network input
        and the scaled residuals of the half- and full-window Euler steps; the solver's
        state advances.  Returns the dataset list of (X, target [N, 2C])."""
        W = self.window()
        raw = []
        for i in range(n_windows):
            U0 = self.s.state.clone()
            self._rhs(U0, self._prim, self._f0)
            prim, f0, grad = self._prim.clone(), self._f0.clone(), self.E.gp.clone()
            sc = self._inc_scale(f0, W, torch.empty(self.M, self.C, device=self.device))
            for _ in range(self.k // 2):
                self.s.step()
            rm = (self.s.state - U0 - 0.5 * W[:, None] * f0).float()
            for _ in range(self.k // 2):
                self.s.step()
            r1 = (self.s.state - U0 - W[:, None] * f0).float()
            raw.append((prim, f0, grad, sc, torch.cat([rm, r1], 1)))
            if log and (i + 1) % max(1, n_windows // 5) == 0:
                log(f"generated {i + 1}/{n_windows} windows")
        self.set_stats_from([r[0] for r in raw])
        data = []
        while raw:
            prim, f0, grad, sc, r = raw.pop(0)
            X = torch.empty(self.N, self.KX, dtype=_BF, device=self.device)
            sc2 = sc[self._cm_long].repeat(1, 2)
            data.append((self._features_from(prim, f0, grad, W, sc, X), (r / sc2).contiguous()))
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
    def _apply(self, U, f0, W, y, col, out):
        nat.check(nat.lib().fvm_mp_apply(nat.ptr(U), nat.ptr(f0), nat.ptr(W), nat.ptr(y[:, col:]), 2 * self.C,
                                         nat.ptr(self._sc), nat.ptr(self.cell_mesh), self.N, self.C, nat.ptr(out),
                                         nat.stream_ptr(self.device)), "fvm_mp_apply")

    def predict(self, U, out_mid=None, out_end=None):
        """Predicted midpoint and end states of the window from U (into the corrector's
        buffers unless given), with f(U) and the window lengths."""
        X, f0, W, sc = self.features(U)
        y = self.net.forward(X, self.nbr)
        Um = self._Um if out_mid is None else out_mid
        U1 = self._U1 if out_end is None else out_end
        self._apply(U, f0, 0.5 * W, y, 0, Um)
        self._apply(U, f0, W, y, self.C, U1)
        return Um, U1, f0, W

    def defect(self, U0, Um, U1, f0, W) -> np.ndarray:
        """Per mesh, max(||d||, ||e||) / ||U1|| of the Simpson and Hermite-midpoint
        defects over the mesh's whole state (host array); per-variable ratios are kept
        in ``last_defect``."""
        self._rhs(Um, self._prim1, self._fm)
        self._rhs(U1, self._prim1, self._f1)
        nat.check(nat.lib().fvm_mp_defect3(nat.ptr(U0), nat.ptr(Um), nat.ptr(U1), nat.ptr(f0), nat.ptr(self._fm),
                                           nat.ptr(self._f1), nat.ptr(W), nat.ptr(self.off), self.M, self.C,
                                           nat.ptr(self._def), nat.ptr(self._part), nat.stream_ptr(self.device)),
                  "fvm_mp_defect3")
        d = self._def.cpu().numpy()
        C = self.C
        u2 = np.maximum(d[:, 2 * C:].sum(1), 1e-300)
        self.last_defect = np.sqrt(d[:, :C] / np.maximum(d[:, 2 * C:], 1e-300))
        return np.sqrt(np.maximum(d[:, :C].sum(1), d[:, C:2 * C].sum(1)) / u2)

    def advance(self, use_net: bool = True) -> bool:
        """One window: predict and check, or (if the check fails or use_net is False)
        k plain RK4 steps.  Returns True when the prediction was accepted."""
        if use_net:
            U0 = self.s.state
            Um, U1, f0, W = self.predict(U0)
            rel = self.defect(U0, Um, U1, f0, W)
            if np.all(rel <= self.tol):
                self.s.set_state(U1)
                self.n_accept += 1
                return True
            self.n_reject += 1
        for _ in range(self.k):
            self.s.step()
        return False

    def run(self, n_windows: int, use_net: bool = True):
        """n_windows windows; returns (seconds, accepted windows)."""
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        acc = sum(self.advance(use_net) for _ in range(n_windows))
        torch.cuda.synchronize()
        return time.perf_counter() - t0, acc
