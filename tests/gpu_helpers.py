"""Shared set-up for the GPU parity tests: the same mesh and state through the product
(HIP, via the C ABI) and through the CPU oracle."""
import os
import time

import numpy as np
import torch

from fvm_solver_amd.config_fvm import ConfigChannel, ConfigChannel3D
from fvm_solver_amd.fvm_equation import FluidConstitution2D, FluidConstitution3D, FVMEquation
from fvm_solver_amd.mesh_utils.fvm_mesh import mesh_from_tuple
from fvm_solver_amd.mesh_utils.mesh_store import channel_bc_tags
from oracle import fvm_oracle as O

DEV = "cuda:0"

_PROGRESS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpurun_out",
                         "gpu_test_progress.log")


def announce(msg):
    """Print a progress line and append it to gpurun_out/gpu_test_progress.log: pytest
    holds a test's output until the test ends, and a long GPU test must show progress
    while it runs."""
    print(f"[gpu-test] {msg}", flush=True)
    try:
        os.makedirs(os.path.dirname(_PROGRESS), exist_ok=True)
        with open(_PROGRESS, "a") as f:
            f.write(f"{time.strftime('%H:%M:%S')} {msg}\n")
    except OSError:
        pass


def oracle_phys(cfg):
    ins, ex = cfg.inlet_cfg, cfg.exit_cfg
    return O.Phys(T_0=cfg.T_0, viscosity=cfg.viscosity, visc_bulk=cfg.visc_bulk, thermal_cond=cfg.thermal_cond,
                  S_const=cfg.S_const, gamma=cfg.gamma, C_v=cfg.C_v, v_factor=cfg.v_factor, lim_p=cfg.lim_p,
                  lim_K=cfg.lim_K,
                  inlet=O.BCSpec(ins.v_n_inf, ins.v_t_inf, ins.rho_inf, ins.T_inf, tuple(ins.v_inf)),
                  exit=O.BCSpec(ex.v_n_inf, ex.v_t_inf, ex.rho_inf, ex.T_inf, tuple(ex.v_inf)))


def make_cfg(dim, **kw):
    cfg = ConfigChannel(device=DEV, **kw) if dim == 2 else ConfigChannel3D(device=DEV, **kw)
    return cfg


def product_equation(pts, cells, fi, fb, tags, cfg, us_init=None):
    dim = cells.shape[1] - 1
    mesh = mesh_from_tuple(pts, cells, fi, fb, device=DEV)
    phy = FluidConstitution2D(cfg) if dim == 2 else FluidConstitution3D(cfg)
    if us_init is None:
        us_init = torch.zeros(mesh.n_cells, dim + 2)
        us_init[:, dim] = 1.0
    eq = FVMEquation(cfg, phy, mesh, dim + 2, channel_bc_tags(tags, dim), us_init=us_init)
    return mesh, eq


def initial_state(centroids, cfg, dim):
    """Freestream plus the Gaussian density bump, as conserved variables (fp32)."""
    c = np.asarray(centroids, np.float64)
    prim = np.zeros((len(c), dim + 2))
    prim[:, 0] = cfg.inlet_cfg.v_n_inf
    r2 = ((c[:, :dim] - np.asarray(cfg.bump_centre[:dim])) ** 2).sum(1)
    prim[:, dim] = 1.0 + cfg.bump_amp * np.exp(-r2 / (2 * cfg.bump_sigma ** 2))
    prim[:, dim + 1] = cfg.inlet_cfg.T_inf
    prim = torch.from_numpy(prim).float()
    V, rho, T = prim[:, :dim], prim[:, dim:dim + 1], prim[:, dim + 1:]
    U = torch.cat([V * rho, rho, rho * (cfg.C_v * T + 0.5 * V.square().sum(-1, keepdim=True))], -1)
    return prim, U


def rel_l2_cols(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b, axis=0) / np.maximum(np.linalg.norm(b, axis=0), 1e-300)


class FluxLedger:
    """Conservation bookkeeping for long runs, per mesh of a device layout and conserved
    variable: the fp64 totals sum_c vol_c U_c, and the trapezoid integral over the steps
    of their rate sum_c vol_c f_c(U) -- interior facet fluxes cancel in that sum, so it is
    the boundary-flux integral -- next to the gross flux scale sum_f |F_f| (what the
    cancellation has to cancel).  f and F come from a separate three-kernel evaluator on
    the same device layout (it keeps its own carried gradients), so the solver is
    untouched.  A step with an interior flux that does not cancel, or a divergence
    weight off, shows as |total change - integral| of the order of the gross scale."""

    def __init__(self, dm, cfg):
        from fvm_solver_amd import _native as nat
        from fvm_solver_amd.fvm_stepping.facet_process import FacetCalc
        self.nat = nat
        self.E = FacetCalc.from_device_mesh(dm, cfg, DEV, fused=False)
        self.dm = dm
        self.C = dm.C
        self.M = dm.n_meshes
        self.vol = dm.volumes64[:, None]
        # cells of a mesh are contiguous, and so are its facets in the device order
        # (sorted by lowest adjacent cell): per-mesh sums are sums over ranges
        self.coff = [int(v) for v in dm.cell_offsets]
        fc = dm.host["fac_cell"][:, dm.host["fperm"]]
        lo = fc.min(0)
        self.foff = [int(v) for v in np.searchsorted(lo, np.asarray(dm.cell_offsets))]
        self.inv_cv = 1.0 / cfg.C_v
        self.prim = torch.empty(dm.N, dm.C, device=DEV)
        self.integral = torch.zeros(self.M, self.C, dtype=torch.float64, device=DEV)
        self.gross = torch.zeros_like(self.integral)
        self.last = None
        self.start = None

    @staticmethod
    def _segsum(x, off):
        return torch.stack([x[off[i]:off[i + 1]].sum(0) for i in range(len(off) - 1)])

    def totals(self, U):
        return self._segsum(self.vol * U.double(), self.coff)

    def _rates(self, U):
        nat = self.nat
        nat.check(nat.lib().fvm_state_to_prim_f32(nat.ptr(U), nat.ptr(self.prim), self.dm.N, self.dm.dim,
                                                  self.inv_cv, nat.stream_ptr(torch.device(DEV))),
                  "fvm_state_to_prim_f32")
        f = self.E.rhs(self.prim)
        return self._segsum(self.vol * f.double(), self.coff), self._segsum(self.E.fflux.abs().double(), self.foff)

    def sample(self, U, dt=None):
        """Record state U (contiguous [N, C] fp32); dt [M] (host floats): the time since
        the previous sample, per mesh."""
        U = U.contiguous()
        r = self._rates(U)
        if self.last is None:
            self.start = self.totals(U)
        else:
            w = torch.as_tensor(np.asarray(dt, np.float64), device=DEV)[:, None] * 0.5
            self.integral += w * (self.last[0] + r[0])
            self.gross += w * (self.last[1] + r[1])
        self.last = r

    def report(self, U):
        """(total change, boundary-flux integral, gross flux integral) as [M, C] numpy."""
        d = (self.totals(U.contiguous()) - self.start).cpu().numpy()
        return d, self.integral.cpu().numpy(), self.gross.cpu().numpy()
