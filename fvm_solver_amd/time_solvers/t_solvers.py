"""State container and time loop -- drop-in for ``time_fvm/time_solvers/t_solvers.py``.

``FVMCells`` (t_solvers.py:15-42) holds the conserved state [N, C] on the GPU;
``TSolver`` (:45-215) keeps the reference's loop: one warm-up step whose only undone
effect is the state (the adaptive dt and the carried gradients keep their new values,
as in the reference), then steps until ``n_iter`` or ``end_t``, with progress printing
and a NaN check on the ``print_i`` cadence instead of a host sync on every step.
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
    def __init__(self, cells: FVMCells, eq, cfg):
        self.cfg = cfg
        self.device = torch.device(cfg.device)
        self.n_steps = cfg.n_iter
        self.end_t = cfg.end_t
        self.cells = cells
        self.eq = eq
        self.print_i = cfg.print_i or 0
        self.dt = torch.tensor(float(cfg.dt if cfg.dt is not None else 0.0), dtype=torch.float32,
                               device=self.device)

    def set_dt(self, dt: float):
        self.dt.fill_(float(dt))

    def _solve(self):
        backup = self.cells.state.clone()
        self._solve_step(self.dt)                 # warm-up (t_solvers.py:88-93)
        self.cells.state = backup
        if self.n_steps is None and self.end_t is None:
            raise ValueError("Either n_iter or end_t must be set in ConfigFVM")
        t = torch.zeros((), dtype=torch.float64, device=self.device)
        i = 0
        st = time.time()
        while True:
            if self.n_steps is not None and i >= self.n_steps:
                break
            if self.end_t is not None and (i % max(self.print_i, 1) == 0) and float(t) >= self.end_t:
                break
            t += self.dt
            self._solve_step(t)
            if self.print_i and i % self.print_i == 0:
                prim = self.cells.get_values()[0]
                if torch.isnan(prim).any():
                    raise ValueError("Nan detected in primatives")
                irl = (time.time() - st) / self.print_i
                print(f"progress: {i = }, t = {float(t):.4g}, dt = {float(self.dt):.3g}, {irl = :.3g}", flush=True)
                st = time.time()
            i += 1

    def solve(self):
        with torch.inference_mode():
            self._solve()

    def _solve_step(self, t):
        self.cells.update_cells(self._step(t))

    @abstractmethod
    def _step(self, t):
        pass

    def _forward_state(self, U):
        prim, _ = self.cells.convert_state_to_value(U)
        return self.eq.forward(prim)

    def _euler_step(self, U):
        return U + self.dt * self._forward_state(U)
