"""Many independent meshes stepped together: one launch per kernel for the whole batch.

Dataset generation in the reference runs one mesh at a time (``run_fvm.py`` +
``ds_saving``).  Here the meshes are concatenated block-diagonally into one device
layout (``DeviceMesh.batch``): cell, facet and boundary indices are offset per mesh, each
mesh keeps its own physical constants (viscosity and conduction follow its Mach number
and obstacle size), its own characteristic-BC group means and its own time step, and a
fused RK stage covers every cell of every mesh in the same two launches.  Fixed-dt
explicit Runge-Kutta only (the adaptive schemes' global norms are per mesh and are not
batched).
"""
from __future__ import annotations

import numpy as np
import torch

from fvm_solver_amd import _native as nat
from fvm_solver_amd.fvm_stepping.facet_process import DeviceMesh, FacetCalc
from fvm_solver_amd.time_solvers.integrators import _tableau

_SHARED = ("gamma", "C_v", "T_0", "S_const", "v_factor", "lim_p", "lim_K")


class BatchedSolver:
    """RK stepping of a batch of meshes.

    problems: list of (mesh, bc_tags, cfg, us_init [N_i, C] conserved state).
    All cfgs must share gamma, C_v, T_0, S_const, v_factor, lim_p and lim_K; viscosity,
    bulk viscosity, conduction, boundary states and dt may differ per mesh."""

    def __init__(self, problems, device="cuda", tableau: str = "RK4", persist_blocks: int = 0):
        cfg0 = problems[0][2]
        for _, _, cfg, _ in problems[1:]:
            for k in _SHARED:
                if getattr(cfg, k) != getattr(cfg0, k):
                    raise ValueError(f"batched meshes must share {k}")
        self.device = torch.device(device)
        dim = problems[0][0].dim
        C = dim + 2
        dm = DeviceMesh.batch([(m, C, tags, cfg) for m, tags, cfg, _ in problems], self.device, persist_blocks)
        self.E = FacetCalc.from_device_mesh(dm, cfg0, self.device)
        self.dm = dm
        self.n_meshes = len(problems)
        self.offsets = dm.cell_offsets
        self.C = C
        U = torch.cat([torch.as_tensor(u, dtype=torch.float32) for _, _, _, u in problems]).to(self.device)
        self.A, self.b, self.b2 = _tableau(tableau)
        s = len(self.b)
        if any(self.A[m][j] != 0 for j in range(s) for m in range(j + 2, s)):
            raise NotImplementedError("batched stepping supports tableaus whose stages use the previous stage only")
        self.stages = s
        n = U.shape
        self._U = [U.contiguous(), torch.empty(n, device=self.device)]
        self._prim = [torch.empty(n, device=self.device), torch.empty(n, device=self.device)]
        self._acc = torch.empty(n, device=self.device)
        self.dt = torch.zeros(self.n_meshes, dtype=torch.float32, device=self.device)
        self._cur = 0
        self._flip = 0
        self._prim_valid = False
        self.version = 0             # bumped by every step and set_state
        self.inv_cv = 1.0 / cfg0.C_v
        self.gR = cfg0.gamma * cfg0.C_v * (cfg0.gamma - 1)
        # first and last mesh of every 256-cell tile of the tile-fused stage
        cm = dm.host["cell_mesh"]
        t0 = np.arange(0, dm.N, 256)
        self._tile_mesh = (cm[t0], cm[np.minimum(t0 + 255, dm.N - 1)])
        self.cell_mesh = torch.from_numpy(cm.astype(np.int64)).to(self.device)

    # ------------------------------------------------------------------ state access
    @property
    def state(self) -> torch.Tensor:
        return self._U[self._cur]

    def state_of(self, i: int) -> torch.Tensor:
        return self.state[self.offsets[i]:self.offsets[i + 1]]

    def set_state(self, U: torch.Tensor):
        self._U[self._cur].copy_(U)
        self._prim_valid = False
        self.version += 1

    def primitives(self) -> torch.Tensor:
        self._ensure_prim()
        return self._prim[self._flip]

    def _ensure_prim(self):
        if self._prim_valid:
            return
        U = self.state
        nat.check(nat.lib().fvm_state_to_prim_f32(nat.ptr(U), nat.ptr(self._prim[self._flip]), U.shape[0],
                                                  self.C - 2, self.inv_cv, nat.stream_ptr(self.device)),
                  "fvm_state_to_prim_f32")
        self.E.invalidate_boundary_means()
        self._prim_valid = True

    # ------------------------------------------------------------------ time step
    def cfl_dt(self, cfl: float) -> np.ndarray:
        """Per-mesh CFL time step from the current state (fp64 reductions per mesh)."""
        prim = self.primitives()
        out = np.zeros(self.n_meshes)
        part = torch.empty(256, dtype=torch.float64, device=self.device)
        res = torch.empty(1, dtype=torch.float64, device=self.device)
        for i in range(self.n_meshes):
            a, b = int(self.offsets[i]), int(self.offsets[i + 1])
            nat.check(nat.lib().fvm_cfl_dt(nat.ptr(prim[a:b]), nat.ptr(self.dm.h64[a:b]), b - a, self.C - 2,
                                           self.gR, cfl, nat.ptr(res), nat.ptr(part), nat.stream_ptr(self.device)),
                      "fvm_cfl_dt")
            out[i] = float(res)
        return out

    def set_dt(self, dts):
        self.dt.copy_(torch.as_tensor(np.asarray(dts, np.float32)))

    def tile_mask(self, meshes: np.ndarray) -> torch.Tensor:
        """int32 ids, ascending, of the 256-cell tiles holding a cell of any mesh with
        meshes[i] true (the tiles a masked step() computes, one workgroup each).  At
        least one mesh must be selected (a step without a list steps every mesh)."""
        if not self.E.fused:
            raise RuntimeError("masked steps need the tile-fused stage")
        sel = np.asarray(meshes, bool)
        if sel.shape != (self.n_meshes,) or not sel.any():
            raise ValueError(f"tile_mask: expected {self.n_meshes} flags with at least one set")
        cs = np.concatenate([[0], np.cumsum(sel)])
        lo, hi = self._tile_mesh
        ids = np.flatnonzero(cs[hi + 1] - cs[lo] > 0).astype(np.int32)
        return torch.from_numpy(ids).to(self.device)

    def step(self, tiles: torch.Tensor | None = None):
        """One fused RK step of every mesh of the batch, or, with a tile list from
        tile_mask(), of the meshes it was made for (the other meshes' cells are left
        stale or computed from stale inputs: overwrite them afterwards)."""
        self._ensure_prim()
        U0 = self._U[self._cur]
        out = self._U[1 - self._cur]
        cur = self._prim[self._flip]
        for i in range(self.stages):
            nxt = self._prim[1 - self._flip]
            st = nat.FvmStage(prim=cur.data_ptr(), U0=U0.data_ptr(), acc=self._acc.data_ptr(), acc2=None,
                              kbuf=None, prim_next=nxt.data_ptr(), U_out=out.data_ptr(), dt=self.dt.data_ptr(),
                              stage=i, last=int(i == self.stages - 1), store_k=0,
                              b=float(self.b[i]), b2=float(self.b2[i]),
                              tile_list=None if tiles is None else tiles.data_ptr(),
                              n_tile_list=0 if tiles is None else int(tiles.numel()))
            if i < self.stages - 1 and self.A[i + 1][i] != 0:
                st.n_inc = 1
                st.inc_idx[0] = i
                st.inc_coef[0] = float(self.A[i + 1][i])
            self.E.stage(st)
            self._flip = 1 - self._flip
            cur = nxt
        self._cur = 1 - self._cur
        self.version += 1

    def capture_graph(self, n_steps: int = 2):
        """Capture n_steps batched steps as one HIP graph (n_steps and n_steps * stages even)."""
        if n_steps % 2 or (n_steps * self.stages) % 2:
            raise ValueError("n_steps and n_steps * stages must be even for a replayable graph")
        self.step()                       # makes the carried state and boundary means current
        torch.cuda.synchronize()
        cur, flip = self._cur, self._flip
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for _ in range(n_steps):
                self.step()
        if (self._cur, self._flip) != (cur, flip):
            raise RuntimeError("captured steps did not return to the starting buffers")
        return g


def random_channel_batch(n_meshes: int, seed: int = 0, n_min: int = 5000, n_max: int = 200_000, device="cuda",
                         progress=None):
    """Dataset-generation batch: per mesh, cell count log-uniform in [n_min, n_max],
    obstacle centre x ~ U(0.3, 0.8) and radius ~ U(0.05, 0.15), freestream Mach
    ~ U(0.1, 0.6); state = freestream + Gaussian density bump.  Returns the problems
    list BatchedSolver takes and the per-mesh parameters (with the generated mesh
    tuple, for checks against other implementations).  progress(i, n_cells), if given,
    is called after each mesh."""
    from fvm_solver_amd.config_fvm import channel_config
    from fvm_solver_amd.mesh_gen.channel import ChannelSpec, channel_mesh_2d, tile_order
    from fvm_solver_amd.mesh_utils.fvm_mesh import mesh_from_tuple
    from fvm_solver_amd.mesh_utils.mesh_store import channel_bc_tags
    rng = np.random.default_rng(seed)
    problems, params = [], []
    start = 0                    # the mesh's first cell in the concatenated batch
    for i in range(n_meshes):
        n = int(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        cx, r, mach = rng.uniform(0.3, 0.8), rng.uniform(0.05, 0.15), rng.uniform(0.1, 0.6)
        pts, cells, (fi, fb), tags = channel_mesh_2d(n, seed=seed * 1000 + i,
                                                     spec=ChannelSpec(centre=(cx, 0.5), radius=r))
        # the batch's 256-cell tiles start at multiples of 256 of the concatenated
        # cells: order this mesh's cells in bisection blocks aligned with them
        cells = cells[tile_order(pts[cells].mean(axis=1), offset=start % 256)]
        start += len(cells)
        mesh = mesh_from_tuple(pts, cells, fi, fb, device=device)
        cfg = channel_config(mach=mach, diameter=2 * r, device=str(device))
        c = mesh.geo64["centroids"]
        u = cfg.inlet_cfg.v_n_inf
        rho = 1.0 + cfg.bump_amp * np.exp(-((c - np.array([cx - 2 * r, 0.5])) ** 2).sum(1) / (2 * cfg.bump_sigma ** 2))
        T = cfg.inlet_cfg.T_inf
        U = np.stack([rho * u, np.zeros_like(rho), rho, rho * (cfg.C_v * T + 0.5 * u * u)], 1)
        problems.append((mesh, channel_bc_tags(tags, 2), cfg, torch.from_numpy(U).float()))
        params.append(dict(cells=mesh.n_cells, centre_x=cx, radius=r, mach=mach,
                           mesh_tuple=(pts, cells, fi, fb, tags)))
        if progress is not None:
            progress(i, mesh.n_cells)
    return problems, params
