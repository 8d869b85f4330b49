"""State container and time loop -- drop-in for ``time_fvm/time_solvers/t_solvers.py``.

``FVMCells`` (t_solvers.py:15-42) holds the conserved state [N, C] on the GPU;
``TSolver`` (:45-215) keeps the reference's loop (:77-164): one warm-up step whose only
undone effect is the state (the adaptive dt and the carried gradients keep their new
values, as in the reference), a save at t = 0, then steps until ``n_iter`` or
``end_t``, with progress printing on the ``print_i`` cadence, a NaN check at every print
and plot time, saves every ``save_t`` (``ds_saving.saving.Saver``, asynchronous) and
``exact_interval`` clamping of the step that would cross a save time.

Host synchronisation: the reference reads the device time every step (its ``t`` is a
tensor compared with Python floats).  Here, for fixed-dt integrators, the simulated time
is mirrored on the host (dt is known there), so the loop never waits for the GPU except
at print, plot and NaN checks; adaptive integrators read dt once per step only when a
save interval, plot interval or ``end_t`` needs the current time.
``cfg.profile`` runs the reference's profiling mode (:179-196): 10 warm-up steps, then
10 steps under ``torch.profiler`` (CPU + GPU activity, roctx-visible step ranges),
a summary table and ``trace.json``.
"""
from __future__ import annotations

import time
from abc import ABC, abstractmethod

import torch


class FVMCells:
    """Conserved state [momentum..., rho, energy] per cell (t_solvers.py:15-42)."""

    def __init__(self, n_cells, n_comp, phys_setup, init_val=None, device="cuda"):
        self.device = torch.device(device)
        self.phys_setup = phys_setup
        if init_val is None:
            self.state = torch.zeros(n_cells, n_comp, device=self.device)
        else:
            assert init_val.shape == (n_cells, n_comp), f"Incorrect us init shape {init_val.shape = }"
            self.state = init_val.to(device=self.device, dtype=torch.float32).contiguous()

    def update_cells(self, state_new):
        self.state = state_new

    def get_values(self):
        return self.phys_setup.state_to_primative(self.state)

    def convert_state_to_value(self, state):
        return self.phys_setup.state_to_primative(state)

    def save(self, name="state.pt"):
        torch.save(self.state, name)

    def load(self, name="state.pt"):
        self.state = torch.load(name, weights_only=True).to(self.device)


class TSolver(ABC):
    adaptive = False

    def __init__(self, cells: FVMCells, eq, cfg):
        self.cfg = cfg
        self.device = torch.device(cfg.device)
        self.n_steps = cfg.n_iter
        self.end_t = cfg.end_t
        self.cells = cells
        self.eq = eq
        self.print_i = cfg.print_i or 0
        self.plot_t = getattr(cfg, "plot_t", None)
        self.plot = getattr(cfg, "plot", False)
        self.save_t = getattr(cfg, "save_t", None)
        self.exact_interval = getattr(cfg, "exact_interval", False)
        self.saver = None
        if self.save_t:
            from fvm_solver_amd.ds_saving.saving import Saver
            self.saver = Saver(self.eq.E_props, save_dir=getattr(cfg, "save_dir", None))
        self.dt = torch.tensor(float(cfg.dt if cfg.dt is not None else 0.0), dtype=torch.float32,
                               device=self.device)
        self._dt_host = float(cfg.dt) if cfg.dt is not None else 0.0

    def set_dt(self, dt: float):
        self.dt.fill_(float(dt))
        self._dt_host = float(dt)

    def _check_nan(self):
        prim = self.cells.get_values()[0]
        if torch.isnan(prim).any():
            raise ValueError("Nan detected in primatives")
        return prim

    def _solve(self):
        backup = self.cells.state.clone()
        self._solve_step(self.dt)                 # warm-up (t_solvers.py:88-93)
        self.cells.state = backup
        if self.n_steps is None and self.end_t is None:
            raise ValueError("Either n_iter or end_t must be set in ConfigFVM")
        if self.saver is not None:
            self.saver.save(torch.tensor(0.0, device=self.device), self.eq.E_props, self.cells.get_values()[0])
        need_t = self.end_t is not None or self.save_t or self.plot_t
        t = 0.0                                   # host mirror of the simulated time
        t_dev = torch.zeros((), dtype=torch.float64, device=self.device)
        next_plot_t = 0.0
        save_count = 1
        next_save_t = self.save_t if self.save_t else None
        i = 0
        st = time.time()
        while True:
            if self.n_steps is not None and i >= self.n_steps:
                break
            if self.end_t is not None and t >= self.end_t:
                break
            dt = float(self.dt) if (self.adaptive and need_t) else self._dt_host
            if self.exact_interval and next_save_t is not None and t + dt > next_save_t:
                saved = self.dt.clone()
                self.dt.fill_(next_save_t - t)
                t_dev += self.dt
                t += next_save_t - t
                self._solve_step(t_dev)
                self.dt.copy_(saved)
            else:
                t_dev += self.dt
                t += dt
                self._solve_step(t_dev)
            if self.print_i and i % self.print_i == 0:
                self._check_nan()
                irl = (time.time() - st) / self.print_i
                print(f"progress: {i = }, t = {float(t_dev):.4g}, dt = {float(self.dt):.3g}, {irl = :.3g}", flush=True)
                st = time.time()
            if self.plot_t and t >= next_plot_t:
                next_plot_t = t + self.plot_t
                self._check_nan()
                if self.plot:
                    print(f"plot: t = {t:.5g} (plotting is not part of this library)", flush=True)
            if next_save_t is not None and t >= next_save_t:
                save_count += 1
                next_save_t = save_count * self.save_t
                self.saver.save(t_dev.clone(), self.eq.E_props, self.cells.get_values()[0])
            i += 1
        if self.saver is not None:
            self.saver.flush()

    def _solve_profile(self):
        """The reference's profiling mode (t_solvers.py:179-196)."""
        import torch.profiler
        for i in range(10):
            self._solve_step(i * self.dt)
        with torch.profiler.profile(activities=[torch.profiler.ProfilerActivity.CPU,
                                                torch.profiler.ProfilerActivity.CUDA],
                                    record_shapes=True, with_stack=True) as prof:
            for i in range(10):
                prof.step()
                torch.cuda.nvtx.range_push(f"step {i}")      # roctx range on ROCm
                self._solve_step(i * self.dt)
                torch.cuda.nvtx.range_pop()
        print(prof.key_averages().table(sort_by="self_cuda_time_total", row_limit=10))
        prof.export_chrome_trace("trace.json")

    def solve(self):
        with torch.inference_mode():
            if getattr(self.cfg, "profile", False):
                self._solve_profile()
            else:
                self._solve()

    def _solve_step(self, t):
        self.cells.update_cells(self._step(t))

    @abstractmethod
    def _step(self, t):
        pass

    def _forward_state(self, U):
        prim, _ = self.cells.convert_state_to_value(U)
        fwd = getattr(self.eq, "forward_device", None)    # the solver's states are in device order
        return fwd(prim) if fwd is not None else self.eq.forward(prim)

    def _euler_step(self, U):
        return U + self.dt * self._forward_state(U)
