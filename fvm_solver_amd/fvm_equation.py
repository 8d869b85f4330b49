"""Navier-Stokes residual -- drop-in for ``time_fvm/fvm_equation.py``.

``FVMEquation(cfg, phy_setup, mesh, n_comp, bc_tag, us_init=None)`` and
``forward(primatives) -> dU/dt`` keep the reference's signatures and state semantics
(fvm_equation.py:371-435): ``forward`` is stateful, the previous call's limited
gradients feed the Neumann-boundary non-orthogonal correction.  The whole residual --
boundary values, LSQ gradient, limiter, face reconstruction, the five facet-flux terms
(advection, pressure, viscosity, heating, KT dissipation) and the divergence -- runs in
the fused HIP kernels of libfvm_hip.so (csrc/rhs.hip).
"""
from __future__ import annotations

import torch

from fvm_solver_amd import _native as nat
from fvm_solver_amd.fvm_stepping.facet_process import FacetCalc
from fvm_solver_amd.time_solvers.integrators import get_solver
from fvm_solver_amd.time_solvers.t_solvers import FVMCells


class FluidConstitution:
    """EOS and state conversions (fvm_equation.py:16-87)."""

    def __init__(self, cfg, dim: int):
        self.device = cfg.device
        self.dim = dim
        self.T_0 = cfg.T_0
        self.mu = cfg.viscosity
        self.mu_b = cfg.visc_bulk
        self.S_const = cfg.S_const
        self.C_v = cfg.C_v
        self.gamma = cfg.gamma
        self.R = self.C_v * (cfg.gamma - 1)

    def state_to_primative(self, state: torch.Tensor):
        """(V, rho, T) from (momentum, rho, E), computed by the HIP kernel."""
        nat.require_device(state, "state")
        state = state.contiguous()
        prim = torch.empty_like(state)
        nat.check(nat.lib().fvm_state_to_prim_f32(nat.ptr(state), nat.ptr(prim), state.shape[0], self.dim,
                                                  1 / self.C_v, nat.stream_ptr(state.device)),
                  "fvm_state_to_prim_f32")
        return prim, state

    def primatives_to_state(self, V, rho, T):
        momentum = V * rho
        Q = rho * (self.C_v * T + 0.5 * V.square().sum(dim=-1, keepdim=True))
        return momentum, rho, Q

    def eos_c(self, rho, T):
        return torch.sqrt(self.gamma * self.R * T)

    def eos_P(self, rho, T):
        return self.R * rho * T

    def eos_T(self, rho, P):
        return P / (self.R * rho)


class FluidConstitution2D(FluidConstitution):
    def __init__(self, cfg, dim: int = 2):
        super().__init__(cfg, dim)
        assert self.dim == 2, "Only for 2D fluids"


class FluidConstitution3D(FluidConstitution):
    def __init__(self, cfg, dim: int = 3):
        super().__init__(cfg, dim)
        assert self.dim == 3, "Only for 3D fluids"


class FVMEquation:
    """``mesh`` is the caller's mesh.  When its cell order is too scattered for the
    tile-fused stage (cfg.reorder "auto": tile-aligned bisection order), or on request
    ("tiles", "morton", "rcm"), the device
    works on the same mesh with its cells renumbered (``cell_perm``: device cell i is mesh
    cell cell_perm[i]); ``forward`` then takes and returns mesh-order arrays, while the
    solver's state (``cells.state``, the saved trajectories and ``E_props``) is in device
    order -- ``state_in_mesh_order()`` maps it back."""

    def __init__(self, cfg, phy_setup: FluidConstitution, mesh, n_comp, bc_tag, us_init=None):
        self.cfg = cfg
        self.phy_setup = phy_setup
        self.device = torch.device(cfg.device)
        self.mesh = mesh
        self.n_comp = n_comp
        self.cell_perm = None
        self._inv_perm = None
        mode = getattr(cfg, "reorder", "auto") or "never"
        if mode not in ("auto", "never", "tiles", "morton", "rcm"):
            raise ValueError(f"unknown cfg.reorder {mode!r} (auto, never, tiles, morton or rcm)")
        E = None
        if mode in ("auto", "never"):
            E = FacetCalc(phy_setup, cfg, mesh, n_comp, bc_tag, device=self.device)
        if mode in ("tiles", "morton", "rcm") or (mode == "auto" and E.mesh.tiles is None and
                                                   E.mesh.tiles_reason not in (None, "disabled")):
            from fvm_solver_amd.mesh_utils.reorder import permuted_mesh
            dev_mesh, perm = permuted_mesh(mesh, "tiles" if mode == "auto" else mode, device=self.device)
            E = FacetCalc(phy_setup, cfg, dev_mesh, n_comp, bc_tag, device=self.device)
            self.cell_perm = torch.from_numpy(perm).to(self.device)
            inv = torch.empty_like(self.cell_perm)
            inv[self.cell_perm] = torch.arange(len(perm), device=self.device)
            self._inv_perm = inv
            if us_init is not None:
                us_init = torch.as_tensor(us_init)[torch.from_numpy(perm)]
        self.E_props = E
        self.cells = FVMCells(mesh.n_cells, n_comp, init_val=us_init, phys_setup=phy_setup, device=self.device)
        self.t_solver = get_solver(self.cells, self, cfg)
        self.E_props.clear_temp()

    def solve(self):
        self.t_solver.solve()

    def forward(self, primatives: torch.Tensor) -> torch.Tensor:
        """dU/dt [N, C] for primitives [N, C] in the mesh's cell order
        (fvm_equation.py:406-426)."""
        if self.cell_perm is None:
            return self.E_props.rhs(primatives)
        return self.E_props.rhs(primatives[self.cell_perm].contiguous())[self._inv_perm]

    def forward_device(self, primatives: torch.Tensor) -> torch.Tensor:
        """dU/dt for primitives in the device's cell order (the solver's own states)."""
        return self.E_props.rhs(primatives)

    def to_mesh_order(self, x: torch.Tensor) -> torch.Tensor:
        """A per-cell array in device order, in the mesh's cell order."""
        return x if self.cell_perm is None else x[self._inv_perm]

    def to_device_order(self, x: torch.Tensor) -> torch.Tensor:
        """A per-cell array in the mesh's cell order, in device order."""
        return x if self.cell_perm is None else x[self.cell_perm]

    def state_in_mesh_order(self) -> torch.Tensor:
        return self.to_mesh_order(self.cells.state)

    # ---- no reference counterpart ----
    def conservation_totals(self, state: torch.Tensor | None = None) -> torch.Tensor:
        """sum_c vol_c U_c per conserved variable, accumulated in fp64 on the GPU (the
        solver's state, or ``state`` given in the mesh's cell order)."""
        state = (self.cells.state if state is None else self.to_device_order(state)).contiguous()
        C = self.n_comp
        out = torch.empty(C, dtype=torch.float64, device=self.device)
        part = torch.empty(256 * C, dtype=torch.float64, device=self.device)
        nat.check(nat.lib().fvm_conservation_f64(nat.ptr(state), nat.ptr(self.E_props.mesh.volumes64),
                                                 state.shape[0], C, nat.ptr(out), nat.ptr(part),
                                                 nat.stream_ptr(self.device)), "fvm_conservation_f64")
        return out

    def cfl_dt(self, prim: torch.Tensor, cfl: float) -> torch.Tensor:
        """dt = cfl * min_c h_c / (|V_c| + c_c), h_c = d vol_c / sum_f |n_f| (fp64, device
        scalar); prim in the mesh's cell order."""
        prim = self.to_device_order(prim).contiguous()
        out = torch.empty(1, dtype=torch.float64, device=self.device)
        part = torch.empty(256, dtype=torch.float64, device=self.device)
        gR = self.phy_setup.gamma * self.phy_setup.R
        nat.check(nat.lib().fvm_cfl_dt(nat.ptr(prim), nat.ptr(self.E_props.mesh.h64), prim.shape[0],
                                       self.mesh.dim, gR, cfl, nat.ptr(out), nat.ptr(part),
                                       nat.stream_ptr(self.device)), "fvm_cfl_dt")
        return out
