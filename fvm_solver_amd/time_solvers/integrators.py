"""Time integrators -- drop-in for ``time_fvm/time_solvers/integrators.py``.

``get_solver(cells, equation, cfg)`` keeps the reference's names (integrators.py:11-27).
``Butcher`` and ``ButcherAdapt`` run every stage as ONE fused pass per RHS: the HIP
stage kernel evaluates dU/dt and applies the tableau update in its epilogue, writing
only the running sums and the next stage's primitives (no k_i round trip for tableaus
whose stages depend on the previous stage only, e.g. classical RK4).  The time step is
a device scalar; ``ButcherAdapt`` updates it on the device
(Adaptive.update_stepsize, integrators.py:42-62), so a step never syncs the host.

Solvers with no runnable reference: ``Adams3_PC``/``Adams4_PC`` and ``Euler`` raise
TypeError in the reference (integrators.py:111,130,199,404); here they run the intended
algorithm (forward(prim) called without the extra arguments).  ``AB3`` is the
build-defined fixed-dt Adams-Bashforth-3 (one RHS per step; history starts as (f0, f0)).
"""
from __future__ import annotations

import numpy as np
import torch

from fvm_solver_amd import _native as nat
from fvm_solver_amd.time_solvers.t_solvers import FVMCells, TSolver


def _tableau(name: str):
    """Butcher tableaus (integrators.py:213-302): (A [s][s], b [s], b2 [s]) as float32."""
    if name == "RK4":
        A = [[0, 0, 0, 0], [0.5, 0, 0, 0], [0, 0.5, 0, 0], [0, 0, 1.0, 0]]
        b, b2 = [1 / 6, 1 / 3, 1 / 3, 1 / 6], [1 / 4] * 4
    elif name == "RK3_SSP4":
        A = [[0, 0, 0, 0], [0.5, 0, 0, 0], [0.5, 0.5, 0, 0], [1 / 6, 1 / 6, 1 / 6, 0]]
        b, b2 = [1 / 6, 1 / 6, 1 / 6, 1 / 2], [1 / 4] * 4
    elif name == "RK3_SSP5":
        p, q, r = 0.37726891511710, 0.16352294089771, 0.14831273384724
        A = [[0] * 5, [p, 0, 0, 0, 0], [p, p, 0, 0, 0], [q, q, q, 0, 0], [0.14904059394856, r, r, 0.34217696850008, 0]]
        b = [0.19707596384481, 0.11780316509765, 0.11709725193772, 0.27015874934251, 0.29786487010104]
        b2 = [1 / 5] * 5
    elif name == "RK3_SSP6":
        p, q, r, s = 0.28422, 0.23071, 0.13416, 0.16528
        A = [[0] * 6, [p, 0, 0, 0, 0, 0], [p, p, 0, 0, 0, 0], [q, q, q, 0, 0, 0], [r, r, r, s, 0, 0], [r, r, r, s, p, 0]]
        b, b2 = [0.17016, 0.17016, 0.10198, 0.12563, 0.21604, 0.21604], [1 / 6] * 6
    elif name == "RK4_SSP5":
        A = [[0] * 5, [0.39175222700392, 0, 0, 0, 0], [0.21766909633821, 0.36841059262959, 0, 0, 0],
             [0.08269208670950, 0.13995850206999, 0.25189177424738, 0, 0],
             [0.06796628370320, 0.11503469844438, 0.20703489864929, 0.54497475021237, 0]]
        b = [0.14681187618661, 0.24848290924556, 0.10425883036650, 0.27443890091960, 0.22600748319395]
        b2 = [1 / 5] * 5
    elif name == "RK4_SSP10":
        A = [[0] * 10] + [[1 / 6] * i + [0] * (10 - i) for i in range(1, 5)] + \
            [[1 / 15] * 5 + [1 / 6] * i + [0] * (5 - i) for i in range(0, 5)]
        b, b2 = [1 / 10] * 10, [1 / 5, 0, 0, 3 / 10, 0, 0, 1 / 5, 0, 3 / 10, 0]
    elif name == "Euler":
        A, b, b2 = [[0]], [1.0], [1.0]
    else:
        raise NotImplementedError("Unknown Butcher tableau")
    return np.asarray(A, np.float32), np.asarray(b, np.float32), np.asarray(b2, np.float32)


class _FusedButcher(TSolver):
    """Shared machinery of the fused explicit RK solvers."""

    adaptive = False

    def __init__(self, cells: FVMCells, equation, name: str, cfg):
        super().__init__(cells, equation, cfg)
        self.A, self.b, self.b2 = _tableau(name)
        self.stages = len(self.b)
        s = self.stages
        # k_j must be stored iff a stage after j + 1 uses it
        self.store = [any(self.A[m][j] != 0 for m in range(j + 2, s)) for j in range(s)]
        n = cells.state.shape
        dev = cells.state.device
        self._prim = [torch.empty(n, device=dev), torch.empty(n, device=dev)]
        self._out = [torch.empty(n, device=dev), torch.empty(n, device=dev)]
        self._acc = torch.empty(n, device=dev)
        self._acc2 = torch.empty(n, device=dev) if self.adaptive else None
        self._kbuf = torch.empty((s,) + tuple(n), device=dev) if any(self.store) else None
        self._last_out = None
        self._flip = 0

    def _stage_prim0(self, U0):
        if self._last_out is not None and U0.data_ptr() == self._last_out.data_ptr():
            return self._prim[self._flip]
        p = self._prim[self._flip]
        self.eq.E_props.invalidate_boundary_means()
        nat.check(nat.lib().fvm_state_to_prim_f32(nat.ptr(U0), nat.ptr(p), U0.shape[0], U0.shape[1] - 2,
                                                  1 / self.cfg.C_v, nat.stream_ptr(U0.device)),
                  "fvm_state_to_prim_f32")
        return p

    def _step(self, t):
        U0 = self.cells.state
        if not U0.is_contiguous() or U0.dtype != torch.float32:
            U0 = U0.contiguous().float()
        nat.require_device(U0, "state")
        E = self.eq.E_props
        out = self._out[1] if (self._last_out is not None and self._last_out.data_ptr() == self._out[0].data_ptr()) \
            else self._out[0]
        cur = self._stage_prim0(U0)
        s = self.stages
        for i in range(s):
            nxt = self._prim[1 - self._flip]
            st = nat.FvmStage(prim=cur.data_ptr(), U0=U0.data_ptr(), acc=self._acc.data_ptr(),
                              acc2=self._acc2.data_ptr() if self._acc2 is not None else None,
                              kbuf=self._kbuf.data_ptr() if self._kbuf is not None else None,
                              prim_next=nxt.data_ptr(), U_out=out.data_ptr(), dt=self.dt.data_ptr(),
                              stage=i, last=int(i == s - 1), store_k=int(self.store[i]),
                              b=float(self.b[i]), b2=float(self.b2[i]))
            if i < s - 1:
                terms = [(j, float(self.A[i + 1][j])) for j in range(i + 1) if self.A[i + 1][j] != 0]
                st.n_inc = len(terms)
                for q, (j, c) in enumerate(terms):
                    st.inc_idx[q] = j
                    st.inc_coef[q] = c
            E.stage(st, prim=cur)
            self._flip = 1 - self._flip
            cur = nxt
        if self.adaptive:
            nat.check(nat.lib().fvm_update_stepsize(E.work, U0.shape[1], E.nblk, *self._adapt, self.dt.data_ptr(),
                                                    nat.stream_ptr(U0.device)), "fvm_update_stepsize")
        self._last_out = out
        return out


class Butcher(_FusedButcher):
    """Fixed-dt explicit Runge-Kutta (integrators.py:349-393)."""

    def capture_graph(self, n_steps: int = 2):
        """Capture ``n_steps`` fused steps as one HIP graph (hipGraph via torch.cuda.CUDAGraph).

        The state ping-pongs between two fixed buffers, so ``n_steps * stages`` and
        ``n_steps`` must be even: a replay then starts and ends in the same buffers.  At
        least one eager step must have run (the carried-gradient flag is then set and the
        kernels are loaded).  Replaying advances ``cells.state`` in place by n_steps.
        """
        if n_steps % 2 or (n_steps * self.stages) % 2:
            raise ValueError("n_steps and n_steps * stages must be even for a replayable graph")
        if not self.eq.E_props.has_prev or self._last_out is None:
            raise RuntimeError("run at least one eager step before capturing")
        start = self.cells.state
        flip0 = self._flip
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for _ in range(n_steps):
                self._solve_step(None)
        if self.cells.state.data_ptr() != start.data_ptr() or self._flip != flip0:
            raise RuntimeError("captured steps did not return to the starting buffers")
        # capture recorded but did not execute the kernels; the state is unchanged
        return g


class ButcherAdapt(_FusedButcher):
    """Explicit RK with the embedded b2 error estimate and in-place dt update
    (integrators.py:305-346); steps are never rejected, as in the reference."""
    adaptive = True

    def __init__(self, cells, equation, name, cfg):
        super().__init__(cells, equation, name, cfg)
        # atol, rtol, mtol, order, alpha(E > 1), alpha(E <= 1)  (integrators.py:318)
        self._adapt = (2e-3, 2e-3, 1e-6, 4.0, 0.8, 0.995)


class _AdaptiveHost:
    """Adaptive.update_stepsize (integrators.py:42-62) on device tensors."""

    def _adapt_init(self, order, rtol, atol, mtol, alphas):
        self.order, self.rtol, self.atol, self.mtol = order, rtol, atol, mtol
        self.alphas = torch.tensor(alphas, device=self.device)

    def update_stepsize(self, dU_high, dU_low, Us):
        E = torch.norm(dU_high - dU_low, dim=0) / (self.atol + self.rtol * torch.norm(dU_high, dim=0)
                                                   + self.mtol * torch.norm(Us, dim=0))
        E = E.mean()
        factor = 0.9 * E ** (-1 / self.order)
        alpha = torch.where(E > 1, self.alphas[0], self.alphas[1])
        self.dt.mul_(torch.addcmul(alpha, (1 - alpha), factor))


class RK3_SSP4(TSolver, _AdaptiveHost):
    """Hand-coded 4-stage SSP RK3 with the (U_b - U_0) error proxy (integrators.py:65-94).
    The RHS is the fused HIP residual; the stage combinations are GPU tensor ops."""
    adaptive = True

    def __init__(self, cells, equation, cfg):
        super().__init__(cells, equation, cfg)
        self._adapt_init(4, 2e-3, 2e-3, 1e-6, (0.8, 0.995))

    def _step(self, t):
        U0 = self.cells.state
        Ua = 1 / 2 * (U0 + self._euler_step(U0))
        Ub = 1 / 2 * (Ua + self._euler_step(Ua))
        Uc = 2 / 3 * U0 + 1 / 6 * (Ub + self._euler_step(Ub))
        U1 = 1 / 2 * (Uc + self._euler_step(Uc))
        self.update_stepsize(U1 - U0, Ub - U0, U0)
        return U1


class Adams3PC(TSolver, _AdaptiveHost):
    """Adams-Bashforth-Moulton 3 predictor-corrector (integrators.py:97-154) with the
    intended forward(prim) calls."""
    adaptive = True

    def __init__(self, cells, equation, cfg):
        super().__init__(cells, equation, cfg)
        self._adapt_init(4, 2e-3, 2e-3, 1e-7, (0.9, 0.99))
        self.hist = None

    def _step(self, t):
        U0 = self.cells.state
        if self.hist is None:
            f = self._forward_state(U0)
            self.hist = (f, f)
        f0 = self._forward_state(U0)
        fm1, fm2 = self.hist
        Ua = U0 + self.dt / 36 * (53 * f0 - 22 * fm1 + 6 * fm2)
        fa = self._forward_state(Ua)
        dU_high = self.dt / 12 * (5 * fa + 8 * f0 - fm1)
        dU_low = self.dt / 2 * (fa + f0)
        self.hist = (f0, fm1)
        self.update_stepsize(dU_high, dU_low, U0)
        return U0 + dU_high


class Adams4PC(Adams3PC):
    """Adams-Bashforth-Moulton 4 predictor-corrector (integrators.py:157-210)."""

    def __init__(self, cells, equation, cfg):
        TSolver.__init__(self, cells, equation, cfg)
        self._adapt_init(4, 1e-3, 1e-3, 1e-6, (0.9, 0.99))
        self.hist = None

    def _step(self, t):
        U0 = self.cells.state
        if self.hist is None:
            f = self._forward_state(U0)
            self.hist = (f, f)
        f0 = self._forward_state(U0)
        fm1, fm2 = self.hist
        Ua = U0 + self.dt / 36 * (53 * f0 - 22 * fm1 + 6 * fm2)
        fa = self._forward_state(Ua)
        dU_high = self.dt / 24 * (9 * fa + 19 * f0 - 5 * fm1 + fm2)
        dU_low = self.dt / 12 * (5 * fa + 8 * f0 - fm1)
        self.hist = (f0, fm1)
        self.update_stepsize(dU_high, dU_low, U0)
        return U0 + dU_high


class AB3(_FusedButcher):
    """Fixed-dt Adams-Bashforth 3, one fused RHS per step (build-defined: the reference
    has no runnable Adams solver):  U1 = U0 + dt/12 (23 f0 - 16 f-1 + 5 f-2), with the
    missing history of the first two steps replaced by f0 (so step 1 is forward Euler).
    The history lives in a ring of three device buffers; nothing is allocated per step."""

    def __init__(self, cells, equation, cfg):
        super().__init__(cells, equation, "Euler", cfg)
        n = cells.state.shape
        self._f = [torch.empty(n, device=cells.state.device) for _ in range(3)]
        self._hist = []          # indices into _f, most recent first

    def reset_history(self):
        self._hist = []

    def _step(self, t):
        U0 = self.cells.state
        if not U0.is_contiguous() or U0.dtype != torch.float32:
            U0 = U0.contiguous().float()
        nat.require_device(U0, "state")
        E = self.eq.E_props
        out = self._out[1] if (self._last_out is not None and self._last_out.data_ptr() == self._out[0].data_ptr()) \
            else self._out[0]
        cur = self._stage_prim0(U0)
        nxt = self._prim[1 - self._flip]
        free = [i for i in range(3) if i not in self._hist[:2]][0]
        st = nat.FvmStage(prim=cur.data_ptr(), U0=U0.data_ptr(), acc=None, prim_next=nxt.data_ptr(),
                          U_out=out.data_ptr(), dt=self.dt.data_ptr(), stage=0, last=1, mode=1,
                          n_hist=min(len(self._hist), 2),
                          f_m1=self._f[self._hist[0]].data_ptr() if len(self._hist) >= 1 else None,
                          f_m2=self._f[self._hist[1]].data_ptr() if len(self._hist) >= 2 else None,
                          f_out=self._f[free].data_ptr())
        E.stage(st, prim=cur)
        self._flip = 1 - self._flip
        self._hist = [free] + self._hist[:1]
        self._last_out = out
        return out


def get_solver(cells: FVMCells, equation, cfg) -> TSolver:
    """Solver factory keyed by cfg.solver_name (integrators.py:11-27)."""
    name, extra = cfg.solver_name, cfg.solver_extra
    if name == "RK3_SSP4":
        return RK3_SSP4(cells, equation, cfg)
    if name == "Adams3_PC":
        return Adams3PC(cells, equation, cfg)
    if name == "Adams4_PC":
        return Adams4PC(cells, equation, cfg)
    if name == "Butcher_adapt":
        return ButcherAdapt(cells, equation, extra, cfg)
    if name == "Butcher":
        return Butcher(cells, equation, extra, cfg)
    if name == "Euler":
        return Butcher(cells, equation, "Euler", cfg)
    if name == "AB3":
        return AB3(cells, equation, cfg)
    raise NotImplementedError("Invalid solver.")
