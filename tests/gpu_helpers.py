"""Shared set-up for the GPU parity tests: the same mesh and state through the product
(HIP, via the C ABI) and through the CPU oracle."""
import numpy as np
import torch

from fvm_solver_amd.config_fvm import ConfigChannel, ConfigChannel3D
from fvm_solver_amd.fvm_equation import FluidConstitution2D, FluidConstitution3D, FVMEquation
from fvm_solver_amd.mesh_utils.fvm_mesh import mesh_from_tuple
from fvm_solver_amd.mesh_utils.mesh_store import channel_bc_tags
from oracle import fvm_oracle as O

DEV = "cuda:0"


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
