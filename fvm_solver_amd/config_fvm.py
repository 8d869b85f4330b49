"""Configuration mirror of ``time_fvm/config_fvm.py`` plus the channel-problem presets.

Field names and meanings follow the reference's ``ConfigFVM`` (config_fvm.py:22-67) and
``ConfigBC`` (:12-19), so a reference configuration object can be passed unchanged.
The presets describe the channel problems this package benchmarks: a 2x1 (x1) box with
a circular (spherical) obstacle of radius 0.1, freestream Mach 0.3 along +x, Re 200 on
the obstacle diameter, gamma 1.4, a Gaussian density bump (amplitude 1e-2, sigma 0.05).
"""
from __future__ import annotations

import math
from abc import ABC
from dataclasses import dataclass
from enum import Enum


class BCMode(Enum):
    Isentropic = "Isentropic"
    Characteristic = "characteristic"
    Farfield = "Farfield"


class ConfigBC(ABC):
    mode: BCMode
    v_n_inf: float
    v_t_inf: float
    v_inf: tuple
    rho_inf: float
    T_inf: float


@dataclass
class ConfigFVM(ABC):
    device: str = "cuda"
    compile: bool = False
    profile: bool = False
    problem_setup: str | None = None
    N_comp: int = 4
    solver_name: str = "Butcher_adapt"
    solver_extra: str = "RK3_SSP4"
    dt: float | None = None
    n_iter: int | None = None
    min_A: float | None = None
    max_A: float | None = None
    lnscale: float | None = None
    plot: bool = False
    plot_t: float | None = None
    save_t: float | None = None
    save_dir: str | None = None
    exact_interval: bool = False
    print_i: int | None = None
    end_t: float | None = None
    T_0: float | None = None
    viscosity: float | None = None
    visc_bulk: float | None = None
    thermal_cond: float | None = None
    S_const: float | None = None
    gamma: float | None = None
    C_v: float | None = None
    v_factor: float = 0.1
    lim_p: int = 4
    lim_K: float = 0.1
    exit_cfg: ConfigBC | None = None
    inlet_cfg: ConfigBC | None = None
    # Cell order on the device (no reference counterpart): "auto" renumbers the cells in
    # tile-aligned bisection order when the mesh's own order is too scattered for the
    # tile-fused stage (e.g. gmsh output); "tiles" / "morton" / "rcm" always renumber;
    # "never" keeps the order.
    reorder: str = "auto"


# ---------------------------------------------------------------- channel presets
GAMMA, C_V, T_INF, RHO_INF, MACH, REYNOLDS, DIAMETER = 1.4, 2.0, 100.0, 1.0, 0.3, 200.0, 0.2
R_GAS = C_V * (GAMMA - 1)
U_INF = MACH * math.sqrt(GAMMA * R_GAS * T_INF)
MU = RHO_INF * U_INF * DIAMETER / REYNOLDS


@dataclass
class ChannelInlet(ConfigBC):
    mode: BCMode = BCMode.Characteristic
    v_n_inf: float = U_INF          # inward-positive, as the reference's inlet config
    v_t_inf: float = 0.0
    v_inf: tuple = (U_INF, 0.0, 0.0)
    rho_inf: float = RHO_INF
    T_inf: float = T_INF


@dataclass
class ChannelFarfield(ConfigBC):
    mode: BCMode = BCMode.Characteristic
    v_n_inf: float = -U_INF         # inward-positive: outflow
    v_t_inf: float = 0.0
    v_inf: tuple = (U_INF, 0.0, 0.0)
    rho_inf: float = RHO_INF
    T_inf: float = T_INF


@dataclass
class ConfigChannel(ConfigFVM):
    """2-D channel: Butcher RK4 at CFL 0.5 (dt set from the initial state)."""
    problem_setup: str | None = "channel"
    N_comp: int = 4
    solver_name: str = "Butcher"
    solver_extra: str = "RK4"
    n_iter: int | None = 1000
    cfl: float = 0.5
    n_cells: int = 20000
    seed: int = 1
    T_0: float | None = 100.0
    viscosity: float | None = MU
    visc_bulk: float | None = MU
    thermal_cond: float | None = MU * GAMMA * C_V / 0.72    # Prandtl 0.72
    S_const: float | None = 110.4
    gamma: float | None = GAMMA
    C_v: float | None = C_V
    bump_amp: float = 1e-2
    bump_sigma: float = 0.05
    bump_centre: tuple = (0.25, 0.5)

    def __post_init__(self):
        if self.exit_cfg is None:
            self.exit_cfg = ChannelFarfield()
        if self.inlet_cfg is None:
            self.inlet_cfg = ChannelInlet()


@dataclass
class ConfigChannel3D(ConfigChannel):
    """3-D channel (2x1x1 box, sphere r 0.1 at (0.5, 0.5, 0.5)): RK4, 5 components."""
    N_comp: int = 5
    n_iter: int | None = 500
    n_cells: int = 4_000_000
    bump_centre: tuple = (0.25, 0.5, 0.5)


def channel_config(mach: float = MACH, diameter: float = DIAMETER, dim: int = 2, **kw) -> ConfigChannel:
    """Channel preset at another freestream Mach number / obstacle diameter: Re 200 on
    the diameter and Prandtl 0.72 fix viscosity and conduction."""
    u = mach * math.sqrt(GAMMA * R_GAS * T_INF)
    mu = RHO_INF * u * diameter / REYNOLDS
    v3 = (u, 0.0, 0.0)
    base = dict(viscosity=mu, visc_bulk=mu, thermal_cond=mu * GAMMA * C_V / 0.72,
                inlet_cfg=ChannelInlet(v_n_inf=u, v_inf=v3), exit_cfg=ChannelFarfield(v_n_inf=-u, v_inf=v3))
    base.update(kw)
    return ConfigChannel(**base) if dim == 2 else ConfigChannel3D(**base)
