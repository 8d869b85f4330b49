"""CPU oracle for the FVM step loop.  TEST INFRASTRUCTURE ONLY.

This module restates the reference algorithm in torch on the CPU so that the HIP
product path can be checked against it.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker
(or as the timed CPU baseline); the product in ``fvm_solver_amd`` never imports it.

Every function takes a ``dtype``: ``torch.float32`` restates the reference's own
precision chain (fp32 geometry, fp64 least-squares inverse, fp32 everything else) and
is the parity target; ``torch.float64`` runs the identical algorithm in double
precision and is the error reference ("FP64 oracle").

Pinning (see tests/test_oracle_*.py):
  * mesh preprocessing, operators, ``csr_to_ell``, the RHS on two successive calls and
    10 steps of each integrator are checked against golden vectors written from the
    reference itself (tests/golden/ref_*.npz);
  * closed forms and invariants: LSQ exactness on linear fields, zero row sums of G,
    sum_f s_f n_f = 0 per cell, global conservation (sum vol*dU/dt equals minus the
    boundary flux), free-stream preservation, the RK4 stability polynomial, and the
    limiter's special cases.
  * ``Adams3_PC``, ``Adams4_PC``, ``Euler`` and ``AB3`` (no runnable reference: the
    reference's Adams and Euler solvers raise TypeError, integrators.py:111,199,404) are
    pinned by the closed-form solution of their linear recurrences on y' = λy (power
    sums over the roots of the characteristic polynomial) and by their order; ``cfl_dt`` and ``conservation_totals`` (no
    reference counterpart) are pinned by closed forms.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

# ======================================================================================
# Configuration mirror (config_fvm.py:6-67)
# ======================================================================================


@dataclass
class BCSpec:
    """Characteristic BC parameters of one boundary region (config_fvm.py:12-19)."""
    v_n_inf: float = 0.0
    v_t_inf: float = 0.0
    rho_inf: float = 1.0
    T_inf: float = 100.0
    v_inf: tuple = (0.0, 0.0, 0.0)


@dataclass
class Phys:
    """Physical/stabilisation parameters consumed on the hot path (config_fvm.py:51-63)."""
    T_0: float = 100.0
    viscosity: float = 1e-3
    visc_bulk: float = 1e-2
    thermal_cond: float = 1e-6
    S_const: float = 110.4
    gamma: float = 1.4
    C_v: float = 2.0
    v_factor: float = 0.1
    lim_p: int = 4
    lim_K: float = 0.1
    inlet: BCSpec = field(default_factory=BCSpec)
    exit: BCSpec = field(default_factory=BCSpec)

    @property
    def R(self):
        return self.C_v * (self.gamma - 1)


# Boundary-condition flags per (facet, component), mesh_store.py:5-59.
DIRICH, NEUMAN, FARFIELD, INLET = 1, 2, 4, 8


def facet_bc_table(tags, dim):
    """BC flags/values per boundary facet and component for the channel problem.

    Mirrors the tag -> Facet mapping of run_fvm.py:264-274 / run_fvm_3d.py:231-239:
    NavierWall = Dirichlet V=0, Neumann dRho/dn = dT/dn = 0; Left = inlet
    (characteristic); Right = farfield (characteristic).
    Returns flags [F_b, C] int, U [F_b, C] float (NaN where unset), dUdn [F_b, C].
    """
    C = dim + 2
    nb = len(tags)
    flags = np.zeros((nb, C), np.int64)
    U = np.full((nb, C), np.nan)
    dUdn = np.full((nb, C), np.nan)
    for i, t in enumerate(tags):
        if t == "NavierWall":
            flags[i, :dim] = DIRICH
            U[i, :dim] = 0.0
            flags[i, dim:] = NEUMAN
            dUdn[i, dim:] = 0.0
        elif t == "Left":
            flags[i, :] = INLET
        elif t == "Right":
            flags[i, :] = FARFIELD
        else:
            raise ValueError(f"Unknown edge tag {t}")
    return flags, U, dUdn


# ======================================================================================
# ELL packing and products (utils/ell_kernel.py, utils/sparse.py)
# ======================================================================================


def csr_to_ell(crow: torch.Tensor, col: torch.Tensor, val: torch.Tensor, K: int | None = None):
    """Restates csr_to_ell (ell_kernel.py:225-256): slot k of row i holds the k-th
    stored entry of the row (CSR order), padding is col 0 / val 0, K = max nnz."""
    M = crow.numel() - 1
    nnz_per_row = crow[1:] - crow[:-1]
    if K is None:
        K = int(nnz_per_row.max().item()) if M > 0 else 0
    if torch.any(nnz_per_row > K):
        raise ValueError("Some rows have more nonzeros than K")
    ell_cols = torch.zeros((M, K), dtype=col.dtype)
    ell_vals = torch.zeros((M, K), dtype=val.dtype)
    for k in range(K):
        rows = torch.nonzero(nnz_per_row > k, as_tuple=False).flatten()
        src = crow[rows] + k
        ell_cols[rows, k] = col[src]
        ell_vals[rows, k] = val[src]
    return ell_vals, ell_cols


def ell_spmm(vals: torch.Tensor, cols: torch.Tensor, x: torch.Tensor, b: torch.Tensor | None = None):
    """y = A_ell @ x (+ b) with accumulation in slot order (ell_kernel.py:24-80,99-132)."""
    x2 = x.unsqueeze(-1) if x.ndim == 1 else x
    M, K = vals.shape
    acc = torch.zeros((M, x2.shape[1]), dtype=x2.dtype)
    for k in range(K):
        acc = acc + vals[:, k:k + 1].to(x2.dtype) * x2[cols[:, k].long()]
    if b is not None:
        acc = acc + b.view(M, 1)
    return acc.squeeze(-1) if x.ndim == 1 else acc


class SPMGeneral:
    """Restates SPMGeneral (sparse.py:41-62): int32 CSR, torch.sparse products."""

    def __init__(self, A: torch.Tensor):
        self.shape = A.shape
        if A.layout != torch.sparse_csr:
            A = A.to_sparse_csr()
        self.A = torch.sparse_csr_tensor(A.crow_indices().to(torch.int32), A.col_indices().to(torch.int32),
                                         A.values(), size=A.size())

    def spMM(self, X, b=None):
        return torch.sparse.mm(self.A, X) if b is None else torch.addmm(b, self.A, X)

    def spMV(self, x, b=None):
        return torch.mv(self.A, x) if b is None else torch.addmv(b, self.A, x)

    def csr(self):
        A = self.A
        return A.crow_indices().long(), A.col_indices().long(), A.values()


def lift_sparse_matrix(A_old: torch.Tensor, n_comp: int):
    """Restates lift_sparse_matrix (sparse.py:231-278)."""
    A_old = A_old.coalesce()
    idx, v = A_old.indices(), A_old.values()
    M, N = A_old.size()
    comp = torch.arange(n_comp)
    rows = (idx[0].unsqueeze(1) * n_comp + comp.unsqueeze(0)).reshape(-1)
    cols = (idx[1].unsqueeze(1) * n_comp + comp.unsqueeze(0)).reshape(-1)
    vals = v.unsqueeze(1).expand(v.numel(), n_comp).reshape(-1)
    return torch.sparse_coo_tensor(torch.stack([rows, cols]), vals, size=(M * n_comp, N * n_comp)).coalesce()


def interleave_sparse_rows(tensors):
    """Restates interleave_sparse_rows (sparse.py:281-325)."""
    k = len(tensors)
    m, n = tensors[0].shape
    inds, vals = [], []
    for i, t in enumerate(tensors):
        t = t.coalesce()
        ii = t.indices().clone()
        ii[0] = k * ii[0] + i
        inds.append(ii)
        vals.append(t.values())
    return torch.sparse_coo_tensor(torch.cat(inds, 1), torch.cat(vals), size=(k * m, n)).coalesce()


def combine_facet_operators(A_main, A_bc, b_bc, bc_mask, n_edges, n_cells, n_comp):
    """Restates the live part of combine_facet_operators (sparse.py:176-228)."""
    main_g = torch.where(~bc_mask)[0]
    bc_g = torch.where(bc_mask)[0]
    mi, mv = A_main._indices(), A_main._values()
    bi, bv = A_bc._indices(), A_bc._values()
    rows_m = main_g[mi[0] // n_comp] * n_comp + mi[0] % n_comp
    rows_b = bc_g[bi[0] // n_comp] * n_comp + bi[0] % n_comp
    idx = torch.stack([torch.cat([rows_m, rows_b]), torch.cat([mi[1], bi[1]])])
    A_all = torch.sparse_coo_tensor(idx, torch.cat([mv, bv]), size=(n_edges * n_comp, n_cells * n_comp),
                                    dtype=A_main.dtype).coalesce()
    b_all = torch.zeros(n_edges * n_comp, dtype=b_bc.dtype)
    r = torch.arange(b_bc.numel())
    b_all[bc_g[r // n_comp] * n_comp + r % n_comp] = b_bc
    return A_all.to_sparse_csr(), b_all


# ======================================================================================
# Mesh preprocessing (mesh_utils/fvm_mesh.py)
# ======================================================================================

FACE_PATTERNS = {2: [(0, 1), (1, 2), (2, 0)], 3: [(0, 1, 2), (0, 2, 3), (0, 3, 1), (1, 3, 2)]}


def _facet_keys(f: np.ndarray, nv: int) -> np.ndarray:
    f = np.sort(f, axis=1).astype(np.int64)
    key = f[:, 0]
    for j in range(1, f.shape[1]):
        key = key * nv + f[:, j]
    return key


class OracleMesh:
    """Restates FVMMesh2D (fvm_mesh.py:112-344) and FVMMesh3D (:347-582), vectorised.

    ``vertices`` are cast to ``dtype`` (the reference casts to fp32 at run_fvm.py:25).
    """

    def __init__(self, vertices, cells, facets, bc_mask, dtype=torch.float32):
        self.dtype = dtype
        V = torch.as_tensor(np.asarray(vertices)).to(dtype)
        cells_np = np.asarray(cells).astype(np.int64)
        facets_np = np.asarray(facets).astype(np.int64)
        self.vertices = V
        self.cells = torch.from_numpy(cells_np)
        self.facets = torch.from_numpy(facets_np)
        self.bc_facet_mask = torch.as_tensor(np.asarray(bc_mask), dtype=torch.bool)
        self.dim = dim = cells_np.shape[1] - 1
        self.n_cells = N = cells_np.shape[0]
        self.n_facets = F = facets_np.shape[0]
        self.n_bc_facet = int(self.bc_facet_mask.sum())
        nl = dim + 1

        # Facet normals and midpoints (fvm_mesh.py:131-139, 367-375, 409-420)
        fv = V[self.facets]
        if dim == 2:
            fvec = fv[:, 1] - fv[:, 0]
            normals = torch.stack([fvec[:, 1], -fvec[:, 0]], dim=1)
        else:
            normals = torch.cross(fv[:, 1] - fv[:, 0], fv[:, 2] - fv[:, 0], dim=1)
        self.normals = normals
        self.midpoints = midpoints = torch.mean(fv, dim=1)

        # Cell volumes and centroids (fvm_mesh.py:141-144, 172-183, 377-380, 422-433)
        cp = V[self.cells]
        if dim == 2:
            ab = cp[:, 1] - cp[:, 0]
            ac = cp[:, 2] - cp[:, 0]
            self.volumes = 0.5 * torch.abs(ab[:, 0] * ac[:, 1] - ab[:, 1] * ac[:, 0])
        else:
            a, b, c, d = cp[:, 0], cp[:, 1], cp[:, 2], cp[:, 3]
            det = torch.sum((b - a) * torch.cross(c - a, d - a, dim=1), dim=1)
            self.volumes = torch.abs(det) / 6.0
        self.centroids = centroids = torch.mean(cp, dim=1)

        # cell -> facet in local face-pattern order (fvm_mesh.py:185-225, 435-473)
        nv = int(V.shape[0])
        fkey = _facet_keys(facets_np, nv)
        order = np.argsort(fkey, kind="stable")
        sk = fkey[order]
        if np.any(sk[1:] == sk[:-1]):
            raise ValueError("duplicate facets")
        ckeys = np.stack([_facet_keys(cells_np[:, list(p)], nv) for p in FACE_PATTERNS[dim]], 1)
        pos = np.searchsorted(sk, ckeys)
        if np.any(pos >= F) or np.any(sk[np.minimum(pos, F - 1)] != ckeys):
            raise KeyError("cell facet missing from the facet list")
        c2f = order[pos]
        self.cell_to_facet = torch.from_numpy(c2f)

        # Signs of each cell's facets (fvm_mesh.py:74-109)
        mid_c = midpoints[self.cell_to_facet]
        nrm_c = normals[self.cell_to_facet]
        p_diff = mid_c - centroids.unsqueeze(1)
        norm_hat = nrm_c / torch.norm(nrm_c, dim=-1, keepdim=True)
        signs = torch.sign(torch.sum(norm_hat * p_diff, dim=-1)).long()
        self.cell_facet_signs = signs
        self.cent_to_facet_disp = p_diff

        # facet -> cells, ordered [s=+1 cell, s=-1 cell] (fvm_mesh.py:93-107, 227-259)
        flat_f = c2f.reshape(-1)
        flat_cell = np.repeat(np.arange(N), nl)
        flat_loc = np.tile(np.arange(nl), N)
        o = np.argsort(flat_f, kind="stable")
        counts = np.bincount(flat_f, minlength=F)
        if np.any(counts == 0) or np.any(counts > 2):
            raise ValueError("facet not shared by one or two cells")
        start = np.concatenate([[0], np.cumsum(counts)[:-1]])
        first_cell = flat_cell[o][start]
        first_loc = flat_loc[o][start]
        two = counts == 2
        second_cell = np.where(two, flat_cell[o][np.minimum(start + 1, len(o) - 1)], -1)
        second_loc = np.where(two, flat_loc[o][np.minimum(start + 1, len(o) - 1)], -1)
        sg = signs.numpy()
        s1 = sg[first_cell, first_loc]
        s2 = np.where(two, sg[np.maximum(second_cell, 0), np.maximum(second_loc, 0)], 0)
        flip = two & (s1 == -1) & (s2 == 1)
        left = np.where(flip, second_cell, first_cell)
        right = np.where(flip, first_cell, second_cell)
        bcm = self.bc_facet_mask.numpy()
        if np.any(two == bcm):
            raise ValueError("Inconsistent boundary bug")
        self.facet_left = left
        self.facet_right = right
        self.facet_to_cell_main = torch.from_numpy(np.stack([left[~bcm], right[~bcm]], 1))
        self.facet_to_cell_bc = torch.from_numpy(left[bcm])

        # Least-squares gradient weights (fvm_mesh.py:263-344, 507-582)
        bc_local = np.full(F, -1, np.int64)
        bc_local[bcm] = np.arange(int(bcm.sum()))
        other = np.where(left[c2f] == np.arange(N)[:, None], right[c2f], left[c2f])
        is_bc = bcm[c2f]
        combined = np.where(is_bc, bc_local[c2f] + N, other)
        self.combined_neigh = torch.from_numpy(combined).int()
        combined_bc = torch.from_numpy(is_bc)
        neigh_cents = torch.where(combined_bc.unsqueeze(-1), midpoints[self.cell_to_facet],
                                  centroids[torch.from_numpy(np.maximum(other, 0))])
        d = neigh_cents - centroids.unsqueeze(1)
        w = 1 / torch.norm(d.double(), dim=2) ** 0.25
        w[combined_bc] *= 7.5
        w2 = w ** 2
        dT = d.transpose(1, 2).double()
        A = torch.bmm(dT, d * w2.unsqueeze(-1))
        A_inv = torch.inverse(A.double())
        G = torch.bmm(A_inv, dT * w2.unsqueeze(1)).to(dtype)      # [N, dim, nl]
        self.G_dense = G
        self.G_mats = []
        rows = torch.arange(N)
        for i in range(dim):
            diag = torch.zeros(N, dtype=dtype)
            for k in range(nl):
                diag = diag + G[:, i, k]
            r = torch.cat([rows.unsqueeze(1).expand(N, nl), rows.unsqueeze(1)], 1).reshape(-1)
            c = torch.cat([self.combined_neigh.long(), rows.unsqueeze(1)], 1).reshape(-1)
            v = torch.cat([G[:, i, :], -diag.unsqueeze(1)], 1).reshape(-1)
            self.G_mats.append(torch.sparse_coo_tensor(torch.stack([r, c]), v, (N, N + self.n_bc_facet)))

        # Centroid displacements (interior) and signed wall distances (boundary)
        lm, rm = left[~bcm], right[~bcm]
        self.cell_disps = centroids[torch.from_numpy(rm)] - centroids[torch.from_numpy(lm)]
        nb_ = normals[self.bc_facet_mask]
        n_hat = nb_ / torch.norm(nb_, dim=-1, keepdim=True)
        fp = midpoints[self.bc_facet_mask] - centroids[self.facet_to_cell_bc]
        dotp = torch.sum(fp * n_hat, dim=-1)
        disp = n_hat * dotp.unsqueeze(-1)
        self.facet_dist_bc = torch.norm(disp, dim=-1) * torch.sign(dotp)


# ======================================================================================
# Operator assembly + RHS (fvm_stepping/*, fvm_equation.py)
# ======================================================================================


class OracleFVM:
    """Restates FacetCalc/MeshCache/BoundarySetter/BC/FVMEquation on the CPU.

    ``forward(prim)`` is FVMEquation.forward (fvm_equation.py:406-426) including its
    state: the limited gradients of the previous call feed the Neumann-boundary
    non-orthogonal correction (facet_process.py:338,346; facet_boundary.py:135-136).
    """

    def __init__(self, mesh: OracleMesh, phys: Phys, bc_flags, bc_U, bc_dUdn):
        self.m = m = mesh
        self.p = p = phys
        dt_ = self.dtype = mesh.dtype
        self.dim = dim = m.dim
        self.C = C = dim + 2
        self.nl = nl = dim + 1
        N, F, Fb = m.n_cells, m.n_facets, m.n_bc_facet
        self.N, self.F, self.Fb = N, F, Fb
        bcm = m.bc_facet_mask
        self.cell_grads = None

        # ---- MeshCache.__init__ (facet_process.py:58-105) ----
        self.slots = (-m.cell_facet_signs + 1 / 2).to(torch.int32).view(nl * N)
        self.cell_to_facet = m.cell_to_facet.view(nl * N)
        self.facet_dists_bc = m.facet_dist_bc.unsqueeze(-1).expand(-1, C)
        self.G = SPMGeneral(interleave_sparse_rows(m.G_mats))
        self.neigh_combine = m.combined_neigh
        self.cent_to_facet_disp = m.cent_to_facet_disp.unsqueeze(-1)
        self.normals = m.normals
        self.facet_size = torch.norm(m.normals, dim=1, keepdim=True)
        n_hat = m.normals / self.facet_size
        self.normals_hat = n_hat.unsqueeze(-1)
        cdf = torch.full((F, dim), float("nan"), dtype=dt_)
        cdf[~bcm] = m.cell_disps
        dcos = (n_hat * cdf).sum(dim=1)
        self.cell_dist_proj = dcos
        X = cdf / dcos.unsqueeze(-1) - n_hat
        X[bcm] = 0
        self.X_orthog = X.unsqueeze(-1)
        assigned = torch.zeros((F, dim), dtype=torch.bool)
        assigned[self.cell_to_facet, self.slots.long()] = True
        self.bc_facet_side = (~assigned[bcm]).to(dt_).argmax(dim=1).int()
        self.bc_locations = torch.where(bcm)[0]
        rows = torch.arange(N).unsqueeze(1).expand(N, nl).reshape(-1)
        vals = self.slots.to(dt_) * 2 - 1
        vinv = (1.0 / m.volumes).to(dt_)
        fm = torch.sparse_coo_tensor(torch.stack([rows, self.cell_to_facet]), vals * vinv[rows], (N, F)).coalesce()
        self.flux_mat = SPMGeneral(fm)

        # ---- SlopeLimiter.__init__ (limiter.py:9-27) ----
        self.eps_p = (p.lim_K * m.volumes.view(-1, 1, 1) ** 0.5) ** (p.lim_p + 1)

        # ---- FacetCalc._init_bc (facet_process.py:297-330) ----
        flags = torch.as_tensor(bc_flags)
        self.dirich_mask = (flags & DIRICH) != 0
        self.neumann_mask = (flags & NEUMAN) != 0
        farfield_mask = torch.all((flags & FARFIELD) != 0, dim=1)
        inlet_mask = torch.all((flags & INLET) != 0, dim=1)
        dv = torch.as_tensor(np.asarray(bc_U), dtype=torch.float32)
        nv = torch.as_tensor(np.asarray(bc_dUdn), dtype=torch.float32)
        self.dirich_val = dv[self.dirich_mask].to(dt_)
        self.neumann_val = nv[self.neumann_mask].to(dt_)
        self._init_boundary(farfield_mask, inlet_mask)
        self._build_face_grads()

    # ---------------------------------------------------------------- BoundarySetter
    def _init_boundary(self, farfield_mask, inlet_mask):
        """Restates BoundarySetter.__init__ and BC.__init__ (facet_boundary.py:32-122, 289-385)
        for the characteristic mode (the only mode the shipped configs use)."""
        m, p, dim, C, N, Fb = self.m, self.p, self.dim, self.C, self.N, self.Fb
        dt_ = self.dtype
        bcm = m.bc_facet_mask
        n_hat_b = self.normals_hat.squeeze(-1)[bcm]
        specs = []
        if bool(farfield_mask.any()):
            sgn = 2 * (self.bc_facet_side[farfield_mask] - 0.5)
            specs.append((p.exit, farfield_mask, m.facet_to_cell_bc[farfield_mask],
                          n_hat_b[farfield_mask] * sgn.unsqueeze(-1).to(dt_)))
        if bool(inlet_mask.any()):
            sgn = 2 * (self.bc_facet_side[inlet_mask] - 0.5)
            specs.append((p.inlet, inlet_mask, m.facet_to_cell_bc[inlet_mask],
                          n_hat_b[inlet_mask] * sgn.unsqueeze(-1).to(dt_)))
        self.groups = []
        if specs:
            merged = specs[0][1].clone()
            for s in specs[1:]:
                merged = merged | s[1]
            bc_idx = torch.where(merged)[0]
            face_to_pos = torch.full((Fb,), -1, dtype=torch.long)
            off = 0
            for s in specs:
                idx = torch.where(s[1])[0]
                face_to_pos[idx] = torch.arange(off, off + idx.shape[0])
                off += idx.shape[0]
            perm = face_to_pos[bc_idx]
            normals = torch.cat([s[3] for s in specs])[perm]
            T_inf = torch.cat([torch.full((int(s[1].sum()),), s[0].T_inf, dtype=dt_) for s in specs])[perm]
            rho_inf = torch.cat([torch.full((int(s[1].sum()),), s[0].rho_inf, dtype=dt_) for s in specs])[perm]
            vts, vns = [], []
            for s in specs:
                n = int(s[1].sum())
                if dim == 2:
                    tang = torch.stack((-s[3][:, 1], s[3][:, 0]), dim=1)
                    vts.append(tang * s[0].v_t_inf)
                    vns.append(torch.full((n,), -s[0].v_n_inf, dtype=dt_))
                else:
                    v_inf = torch.tensor(s[0].v_inf, dtype=dt_)
                    vn = (s[3] * v_inf).sum(dim=1)
                    vts.append(v_inf - s[3] * vn.unsqueeze(-1))
                    vns.append(vn)
            v_t_inf = torch.cat(vts)[perm]
            v_n_inf = torch.cat(vns)[perm]
            p_inf = torch.cat([self.eos_P(torch.full((int(s[1].sum()),), s[0].rho_inf, dtype=dt_),
                                          torch.full((int(s[1].sum()),), s[0].T_inf, dtype=dt_)) for s in specs])[perm]
            cells_g = torch.cat([s[2] for s in specs])[perm]
            self.groups.append(dict(bc_idx=bc_idx, cells=cells_g, normals=normals, T_inf=T_inf, rho_inf=rho_inf,
                                    v_t_inf=v_t_inf, v_n_inf=v_n_inf, p_inf=p_inf))

        # Neumann non-orthogonal correction geometry (facet_boundary.py:82-116)
        c2f = m.cell_to_facet
        neum_all = torch.zeros((self.F, C), dtype=torch.bool)
        neum_all[bcm] = self.neumann_mask
        wf, wc = torch.where(neum_all)
        where_neum = torch.where(self.neumann_mask)
        bc_f2c = torch.zeros(self.F, dtype=torch.long)
        bc_f2c[bcm] = m.facet_to_cell_bc
        neum_cells = bc_f2c[wf]
        cfid = torch.where(wf.unsqueeze(-1).repeat(1, self.nl) == c2f[neum_cells])[1]
        n_hats = self.normals_hat.squeeze(-1)[wf]
        r = self.cent_to_facet_disp.squeeze(-1)[neum_cells][torch.arange(neum_cells.shape[0]), cfid]
        dn = n_hats * (r * n_hats).sum(dim=1, keepdim=True)
        l_vec = r - dn

        # A_bc, b_bc (facet_boundary.py:238-280); undefined rows are NaN (torch.empty there)
        bc_rows = torch.arange(Fb).unsqueeze(1).expand(Fb, C).reshape(-1)
        comp_idx = torch.arange(C).unsqueeze(0).expand(Fb, C).reshape(-1)
        dm = self.dirich_mask.reshape(-1)
        nm = self.neumann_mask.reshape(-1)
        ni = torch.nonzero(nm).squeeze(1)
        A_cols = m.facet_to_cell_bc[bc_rows[ni]] * C + comp_idx[ni]
        A_bc = torch.sparse_coo_tensor(torch.stack([ni, A_cols]), torch.ones(ni.numel(), dtype=dt_), (Fb * C, N * C))
        self.A_bc = SPMGeneral(A_bc)
        b_bc = torch.full((Fb * C,), float("nan"), dtype=dt_)
        b_bc[dm] = self.dirich_val
        b_bc[nm] = self.neumann_val[comp_idx[nm]] * self.facet_dists_bc.flatten()[nm]
        self.b_bc = b_bc

        # A_corr (facet_boundary.py:170-236)
        nn = neum_cells.shape[0]
        row_ids = where_neum[0] * C + where_neum[1]
        rows = row_ids[:, None].expand(nn, dim).reshape(-1)
        cols = (neum_cells[:, None] * (dim * C) + torch.arange(dim)[None, :] * C
                + wc[:, None].expand(nn, dim)).reshape(-1)
        A_corr = torch.sparse_coo_tensor(torch.stack([rows, cols]), l_vec.reshape(-1), (Fb * C, N * dim * C)).coalesce()
        self.A_corr = SPMGeneral(A_corr)

    def _build_face_grads(self):
        """Restates MeshCache.build_spm_face_grads (facet_process.py:114-198)."""
        m, C, N, Fb = self.m, self.C, self.N, self.Fb
        dt_ = self.dtype
        bcm = m.bc_facet_mask
        nfm = m.facet_to_cell_main.shape[0]
        rows = torch.arange(nfm).repeat_interleave(2)
        cols = m.facet_to_cell_main.reshape(-1)
        dist_rep = self.cell_dist_proj[~bcm].repeat_interleave(2)
        signs = torch.tensor([-1, 1], dtype=dt_).repeat(nfm)
        A_face = torch.sparse_coo_tensor(torch.stack([rows, cols]), signs / dist_rep, size=(nfm, N))
        A_main = lift_sparse_matrix(A_face, C)
        bc_rows = torch.arange(Fb).unsqueeze(1).expand(Fb, C).reshape(-1)
        comp_idx = torch.arange(C).unsqueeze(0).expand(Fb, C).reshape(-1)
        dr = torch.nonzero(self.dirich_mask.reshape(-1)).squeeze(1)
        dfi = bc_rows[dr]
        nr = torch.nonzero(self.neumann_mask.reshape(-1)).squeeze(1)
        ncomp = comp_idx[nr]
        cols = m.facet_to_cell_bc[dfi] * C + comp_idx[dr]
        vals = -1.0 / self.facet_dists_bc[dfi, 0]
        A_gbc = torch.sparse_coo_tensor(torch.stack([dr, cols]), vals, size=(Fb * C, N * C))
        b_grad = torch.zeros(Fb * C, dtype=dt_)
        b_grad[dr] = self.dirich_val / self.facet_dists_bc[dfi, 0]
        b_grad[nr] = self.neumann_val[ncomp]
        A_all, self.b_face_grad = combine_facet_operators(A_main, A_gbc, b_grad, bcm, self.F, N, C)
        self.A_face_grad = SPMGeneral(A_all)

    # ---------------------------------------------------------------- EOS helpers
    def eos_P(self, rho, T):
        return self.p.R * rho * T

    def eos_c(self, rho, T):
        return torch.sqrt(self.p.gamma * self.p.R * T)

    def eos_T(self, rho, P):
        return P / (self.p.R * rho)

    def state_to_primative(self, state):
        """fvm_equation.py:35-44"""
        d = self.dim
        mom, rho, Q = state[:, :d], state[:, d:d + 1], state[:, d + 1:d + 2]
        V = mom / rho
        T = 1 / self.p.C_v * (Q / rho - 0.5 * V.square().sum(dim=1, keepdim=True))
        return torch.cat([V, rho, T], dim=-1)

    def primatives_to_state(self, V, rho, T):
        """fvm_equation.py:46-55"""
        return V * rho, rho, rho * (self.p.C_v * T + 0.5 * V.square().sum(dim=-1, keepdim=True))

    # ---------------------------------------------------------------- boundary values
    def set_face_values(self, Us, cell_grads=None):
        """Restates BoundarySetter.set_face_values + BC.BC_characteristic
        (facet_boundary.py:124-168, 418-512)."""
        dim, C = self.dim, self.C
        U_flat = self.A_bc.spMV(Us.flatten(), self.b_bc)
        if cell_grads is not None:
            U_flat.add_(self.A_corr.spMV(cell_grads.reshape(-1)))
        U_face = U_flat.view(self.Fb, C)
        for g in self.groups:
            Ub = Us[g["cells"]]
            n = g["normals"]
            Vs, rho_i, T_i = Ub[:, :dim], Ub[:, dim], Ub[:, dim + 1]
            v_n = (Vs * n).sum(dim=1)
            v_t = Vs - v_n.unsqueeze(-1) * n
            p_i = self.eos_P(rho_i, T_i)
            c_i = self.eos_c(rho_i, T_i)
            drho = g["rho_inf"] - rho_i
            dvn = g["v_n_inf"] - v_n
            dp = g["p_inf"] - p_i
            inv_c = 1 / c_i
            inv_c2 = inv_c ** 2
            h_dp = 0.5 * inv_c2 * dp
            h_dvn = 0.5 * rho_i * inv_c * dvn
            x0 = -h_dvn + h_dp
            x1 = drho - inv_c2 * dp
            x2 = h_dvn + h_dp
            scale = 10.0 / c_i.mean()
            g0 = 0.5 * (1.0 - torch.tanh(scale * (v_n - c_i)))
            g1 = 0.5 * (1.0 - torch.tanh(scale * v_n))
            g2 = 0.5 * (1.0 - torch.tanh(scale * (v_n + c_i)))
            x0, x1, x2 = g0 * x0, g1 * x1, g2 * x2
            g1 = g1.unsqueeze(1)
            v_t_b = g1 * v_t + (1.0 - g1) * g["v_t_inf"]
            c2 = c_i * c_i
            cor = c_i / rho_i
            rho_b = rho_i + (x0 + x1 + x2)
            v_n_b = v_n + (-cor * x0 + cor * x2)
            p_b = p_i + c2 * (x0 + x2)
            V = v_n_b.unsqueeze(-1) * n + v_t_b
            T_b = self.eos_T(rho_b, p_b)
            U_face[g["bc_idx"]] = torch.cat((V, rho_b.unsqueeze(-1), T_b.unsqueeze(-1)), dim=-1)
        return U_face

    # ---------------------------------------------------------------- limiter
    def _limit(self, delta, dU):
        """SlopeLimiter.p1-p5 + limit (limiter.py:29-87)."""
        lp, eps = self.p.lim_p, self.eps_p
        if lp == 1:
            dU = 2 * ((dU > 0).to(delta.dtype) - 0.5) * (dU.abs() + 1e-8)
            phi = torch.clamp(delta / dU, min=0.0, max=1)
        elif lp == 2:
            phi = (delta ** 2 + eps + 2 * delta * dU) / (delta ** 2 + 2 * dU ** 2 + delta * dU + eps)
        elif lp == 3:
            a, b = delta.abs(), dU.abs()
            a_eps = a ** 3 + eps
            S = 4 * b ** 2
            phi = torch.where(a < 2 * b, (a_eps + a * S) / (a_eps + b * (delta ** 2 + S)), 1)
        elif lp == 4:
            a, b = delta.abs(), dU.abs()
            b2 = 2 * b
            a_eps = a ** 4 + eps
            S = b2 * torch.addcmul(a.square(), b2, b2 - a)
            phi = torch.addcmul(a_eps, a, S) / torch.addcmul(a_eps, b, (delta ** 3 + S))
            phi = torch.where(a < b2, phi, 1)
        elif lp == 5:
            a, b = delta.abs(), dU.abs()
            a_eps = a ** 5 + eps
            S = 8 * b ** 2 * (a ** 2 - 2 * b * (a - b))
            phi = torch.where(a < 2 * b, (a_eps + a * S) / (a_eps + b * (delta ** 4 + S)), 1)
        else:
            raise NotImplementedError(f"Limiter of order {lp} is not implemented.")
        return torch.amin(phi, dim=1, keepdim=True)

    # ---------------------------------------------------------------- RHS
    def forward(self, Us):
        """FVMEquation.forward (fvm_equation.py:406-435) with FacetCalc.compute_facet
        (facet_process.py:332-450); returns dU/dt [N, C]."""
        dim, C, N, nl, F = self.dim, self.C, self.N, self.nl, self.F
        dt_ = self.dtype
        p = self.p
        U_bc = self.set_face_values(Us, self.cell_grads)
        Ucf = torch.cat([Us, U_bc])
        grads = self.G.spMM(Ucf).view(N, dim, C)
        # _limit_face_vals (facet_process.py:387-407)
        Uc = Us.unsqueeze(1)
        Un = Ucf[self.neigh_combine.long()]
        dU = (grads.unsqueeze(1) * self.cent_to_facet_disp).sum(dim=2)
        Ucn = torch.cat([Uc, Un], dim=1)
        upper = torch.amax(Ucn, dim=1, keepdim=True) - Uc
        lower = torch.amin(Ucn, dim=1, keepdim=True) - Uc
        phi_lim = self._limit(torch.where(dU > 0, upper, lower), dU)
        Us_face = torch.addcmul(Uc, dU, phi_lim).view(nl * N, C)
        grads = grads * phi_lim
        self.cell_grads = grads
        self.last_ubc = U_bc
        Ufa = torch.empty((F, 2, C), dtype=dt_)
        Ufa[self.cell_to_facet, self.slots.long()] = Us_face
        Ufa[self.bc_locations, self.bc_facet_side.long()] = U_bc
        self.last_face = Ufa                 # the reference's E_props.*_facet (Saver input)
        Vs = Ufa[:, :, :dim].contiguous()
        rho = Ufa[:, :, dim].unsqueeze(-1).contiguous()
        T = Ufa[:, :, dim + 1].unsqueeze(-1).contiguous()
        mom, _, Q = self.primatives_to_state(Vs, rho, T)
        phi = (Vs * self.normals.unsqueeze(1)).sum(dim=-1)
        gc = list(range(dim)) + [dim + 1]
        gF = grads[:, :, gc].reshape(N, dim * (dim + 1))
        gF_flat = torch.repeat_interleave(gF, nl, dim=0, output_size=nl * N)
        gfa = torch.empty((F, 2, dim * (dim + 1)), dtype=dt_)
        gfa[self.cell_to_facet, self.slots.long()] = gF_flat
        gfa[self.bc_locations, self.bc_facet_side.long()] = gF[self.m.facet_to_cell_bc]
        gls = gfa.mean(dim=1).reshape(F, dim, dim + 1)
        gdn = self.A_face_grad.spMV(Us.flatten(), self.b_face_grad).view(F, C)[:, gc]
        dFdn = gdn - (gls * self.X_orthog).sum(dim=1)
        gdotn = (gls * self.normals_hat).sum(dim=1, keepdim=True)
        gFc = gls + (dFdn.unsqueeze(1) - gdotn) * self.normals_hat
        grad_V = gFc[:, :, :dim]
        grad_T_n = dFdn[:, -1]
        self.Vs_facet, self.rho_facet, self.T_facet = Vs, rho, T
        self.grad_V, self.grad_T_n = grad_V, grad_T_n

        # FluidConstitution{2D,3D}._tau / _pressure (fvm_equation.py:62-187)
        Tm = T.mean(dim=1).squeeze()
        D = 0.5 * (grad_V + grad_V.permute(0, 2, 1))
        I1 = D[:, 0, 0] + D[:, 1, 1]
        mu = p.viscosity * (Tm / p.T_0) ** 1.5 * (p.T_0 + p.S_const) / (Tm + p.S_const)
        mu_b = p.visc_bulk * Tm ** 2 / p.T_0 ** 2
        a0 = (-mu_b * I1).view(-1, 1, 1)
        a1 = (-2 * mu).view(-1, 1, 1)
        eye = torch.eye(dim, dtype=dt_)
        tau = a0 * eye + a1 * D if dim == 2 else a0 * eye + a1 * D + 0 * (D @ D)
        P = self.eos_P(rho, T)
        c = self.eos_c(rho, T)

        # Advection, pressure, viscosity, heating, KT dissipation (fvm_equation.py:210-368)
        Usf = torch.cat([mom, rho, Q + P], dim=-1)
        fl = (Usf * phi.unsqueeze(-1)).mean(dim=1).contiguous()
        fl[:, :dim] += P.mean(dim=1) * self.normals
        fl[:, :dim] += (tau * self.normals.unsqueeze(-1)).sum(dim=-2)
        heat = (tau * Vs.mean(dim=1, keepdim=True) * self.normals.unsqueeze(-1)).sum(dim=(-1, -2))
        heat.addcmul_(grad_T_n, self.facet_size.squeeze(), value=-p.thermal_cond)
        fl[:, dim + 1] += heat
        Vn = Vs.norm(dim=-1)
        Vmax = Vn.max(dim=1, keepdim=True).values
        cm = c.max(dim=1).values
        M = (Vmax / (cm + 1e-8)).abs()
        vf = torch.clamp(M, min=p.v_factor, max=1)
        a_vel = vf * cm + Vmax
        a_oth = cm + Vmax
        kt = 0.5 * torch.cat([a_vel * (mom[:, 0] - mom[:, 1]), a_oth * (rho[:, 0] - rho[:, 1]),
                              a_oth * (Q[:, 0] - Q[:, 1])], dim=1) * self.facet_size
        fl += kt
        self.fluxes = fl
        return self.flux_mat.spMM(fl)

    # ---------------------------------------------------------------- build-only checks
    def conservation_totals(self, state):
        """sum_c vol_c * U_c per conserved variable in FP64 (no reference counterpart)."""
        return (self.m.volumes.double().unsqueeze(1) * state.double()).sum(dim=0)

    def cfl_dt(self, prim, cfl):
        """dt = CFL * min_c h_c / (|V_c| + c_c), h_c = d * vol_c / sum_f |n_f| (build-defined)."""
        d = self.dim
        V = prim[:, :d].double()
        c = torch.sqrt(self.p.gamma * self.p.R * prim[:, d + 1].double())
        per = self.facet_size.squeeze(-1).double()[self.m.cell_to_facet].sum(dim=1)
        h = d * self.m.volumes.double() / per
        return float(cfl * torch.min(h / (V.norm(dim=1) + c)))


# ======================================================================================
# Time integration (time_solvers/integrators.py)
# ======================================================================================

def butcher_table(name: str):
    """Butcher_Tables (integrators.py:213-302): returns (A [s,s], b [s], b2 [s]) fp32."""
    f = torch.float32
    if name == "RK4":
        A = torch.tensor([[0, 0, 0, 0], [0.5, 0, 0, 0], [0, 0.5, 0, 0], [0, 0, 1.0, 0]], dtype=f)
        b = torch.tensor([1 / 6, 1 / 3, 1 / 3, 1 / 6], dtype=f)
        b2 = torch.tensor([1 / 4] * 4, dtype=f)
    elif name == "RK3_SSP4":
        A = torch.tensor([[0, 0, 0, 0], [0.5, 0, 0, 0], [0.5, 0.5, 0, 0], [1 / 6, 1 / 6, 1 / 6, 0]], dtype=f)
        b = torch.tensor([1 / 6, 1 / 6, 1 / 6, 1 / 2], dtype=f)
        b2 = torch.tensor([1 / 4] * 4, dtype=f)
    elif name == "RK3_SSP5":
        a1, a2, a3 = 0.37726891511710, 0.16352294089771, 0.14831273384724
        A = torch.tensor([[0] * 5, [a1, 0, 0, 0, 0], [a1, a1, 0, 0, 0], [a2, a2, a2, 0, 0],
                          [0.14904059394856, a3, a3, 0.34217696850008, 0]], dtype=f)
        b = torch.tensor([0.19707596384481, 0.11780316509765, 0.11709725193772, 0.27015874934251,
                          0.29786487010104], dtype=f)
        b2 = torch.tensor([1 / 5] * 5, dtype=f)
    elif name == "RK3_SSP6":
        a, c_, d_, e = 0.28422, 0.23071, 0.13416, 0.16528
        A = torch.tensor([[0] * 6, [a, 0, 0, 0, 0, 0], [a, a, 0, 0, 0, 0], [c_, c_, c_, 0, 0, 0],
                          [d_, d_, d_, e, 0, 0], [d_, d_, d_, e, a, 0]], dtype=f)
        b = torch.tensor([0.17016, 0.17016, 0.10198, 0.12563, 0.21604, 0.21604], dtype=f)
        b2 = torch.tensor([1 / 6] * 6, dtype=f)
    elif name == "RK4_SSP5":
        A = torch.tensor([[0] * 5, [0.39175222700392, 0, 0, 0, 0], [0.21766909633821, 0.36841059262959, 0, 0, 0],
                          [0.08269208670950, 0.13995850206999, 0.25189177424738, 0, 0],
                          [0.06796628370320, 0.11503469844438, 0.20703489864929, 0.54497475021237, 0]], dtype=f)
        b = torch.tensor([0.14681187618661, 0.24848290924556, 0.10425883036650, 0.27443890091960,
                          0.22600748319395], dtype=f)
        b2 = torch.tensor([1 / 5] * 5, dtype=f)
    elif name == "RK4_SSP10":
        rows = [[0] * 10, [1 / 6] + [0] * 9, [1 / 6] * 2 + [0] * 8, [1 / 6] * 3 + [0] * 7, [1 / 6] * 4 + [0] * 6,
                [1 / 15] * 5 + [0] * 5, [1 / 15] * 5 + [1 / 6] + [0] * 4, [1 / 15] * 5 + [1 / 6] * 2 + [0] * 3,
                [1 / 15] * 5 + [1 / 6] * 3 + [0] * 2, [1 / 15] * 5 + [1 / 6] * 4 + [0]]
        A = torch.tensor(rows, dtype=f)
        b = torch.tensor([1 / 10] * 10, dtype=f)
        b2 = torch.tensor([1 / 5, 0, 0, 3 / 10, 0, 0, 1 / 5, 0, 3 / 10, 0], dtype=f)
    else:
        raise NotImplementedError("Unknown Butcher tableau")
    return A, b, b2


class OracleStepper:
    """Restates Butcher/ButcherAdapt/RK3_SSP4 (integrators.py:65-94, 305-393) and the
    build-defined Adams schemes.  ``rhs(U)`` converts state -> primitives and calls
    the oracle's forward, exactly as TSolver._forward_state (t_solvers.py:213-215)."""

    def __init__(self, fvm: OracleFVM, name: str, extra: str = "", dt: float = 1e-4):
        self.fvm = fvm
        self.name = name
        self.extra = extra
        dt_ = fvm.dtype
        self.dt = torch.tensor(dt, dtype=torch.float32).to(dt_)
        if name in ("Butcher", "Butcher_adapt"):
            A, b, b2 = butcher_table(extra)
            self.A, self.b, self.b2 = A.to(dt_), b.to(dt_).reshape(-1, 1, 1), b2.to(dt_).reshape(-1, 1, 1)
        if name in ("Butcher_adapt", "RK3_SSP4"):
            self.ad = dict(order=4, atol=2e-3, rtol=2e-3, mtol=1e-6, alphas=torch.tensor((0.8, 0.995)).to(dt_))
        if name == "Adams3_PC":
            self.ad = dict(order=4, atol=2e-3, rtol=2e-3, mtol=1e-7, alphas=torch.tensor((0.9, 0.99)).to(dt_))
        if name == "Adams4_PC":
            self.ad = dict(order=4, atol=1e-3, rtol=1e-3, mtol=1e-6, alphas=torch.tensor((0.9, 0.99)).to(dt_))
        self.hist = None

    def rhs(self, U):
        return self.fvm.forward(self.fvm.state_to_primative(U))

    def _update_stepsize(self, dU_high, dU_low, Us):
        """Adaptive.update_stepsize (integrators.py:42-62)."""
        a = self.ad
        E = torch.norm(dU_high - dU_low, dim=0) / (a["atol"] + a["rtol"] * torch.norm(dU_high, dim=0)
                                                   + a["mtol"] * torch.norm(Us, dim=0))
        E = E.mean()
        factor = 0.9 * E ** (-1 / a["order"])
        alpha = torch.where(E > 1, a["alphas"][0], a["alphas"][1])
        self.dt.mul_(torch.addcmul(alpha, (1 - alpha), factor))

    def step(self, U0):
        dt = self.dt
        if self.name == "Butcher":
            s = self.b.shape[0]
            k = torch.zeros((s, *U0.shape), dtype=U0.dtype)
            for i in range(s):
                inc = 0 if i == 0 else (self.A[i, :i].unsqueeze(-1).unsqueeze(-1) * k[:i]).sum(dim=0)
                k[i] = dt * self.rhs(U0 + inc)
            return U0 + torch.sum(self.b * k, dim=0)
        if self.name == "Butcher_adapt":
            s = self.b.shape[0]
            k = torch.zeros((s, *U0.shape), dtype=U0.dtype)
            k[0] = dt * self.rhs(U0)
            A = self.A.unsqueeze(-1).unsqueeze(-1)
            for i in range(1, s):
                inc = torch.einsum("j...,j...->...", A[i, :i], k[:i])
                k[i] = dt * self.rhs(U0 + inc)
            dU_high = torch.sum(self.b * k, dim=0)
            dU_low = torch.sum(self.b2 * k, dim=0)
            self._update_stepsize(dU_high, dU_low, U0)
            return U0 + dU_high
        if self.name == "RK3_SSP4":
            def euler(U):
                return U + self.dt * self.rhs(U)
            Ua = 1 / 2 * (U0 + euler(U0))
            Ub = 1 / 2 * (Ua + euler(Ua))
            Uc = 2 / 3 * U0 + 1 / 6 * (Ub + euler(Ub))
            U1 = 1 / 2 * (Uc + euler(Uc))
            self._update_stepsize(U1 - U0, Ub - U0, U0)
            return U1
        if self.name == "Adams3_PC":
            # Adams3PC._step (integrators.py:97-154) with forward(prim) called without the
            # extra arguments that make the reference raise TypeError.
            if self.hist is None:
                f0 = self.rhs(U0)
                self.hist = (f0, f0)
            f0 = self.rhs(U0)
            fm1, fm2 = self.hist
            Ua = U0 + dt / 36 * (53 * f0 - 22 * fm1 + 6 * fm2)
            fa = self.rhs(Ua)
            dU_high = dt / 12 * (5 * fa + 8 * f0 - fm1)
            dU_low = dt / 2 * (fa + f0)
            self.hist = (f0, fm1)
            self._update_stepsize(dU_high, dU_low, U0)
            return U0 + dU_high
        if self.name == "Adams4_PC":
            # Adams4PC._step (integrators.py:157-210): the same predictor as Adams3PC, a
            # four-term corrector and the three-term corrector as the low-order estimate
            if self.hist is None:
                f0 = self.rhs(U0)
                self.hist = (f0, f0)
            f0 = self.rhs(U0)
            fm1, fm2 = self.hist
            Ua = U0 + dt / 36 * (53 * f0 - 22 * fm1 + 6 * fm2)
            fa = self.rhs(Ua)
            dU_high = dt / 24 * (9 * fa + 19 * f0 - 5 * fm1 + fm2)
            dU_low = dt / 12 * (5 * fa + 8 * f0 - fm1)
            self.hist = (f0, fm1)
            self._update_stepsize(dU_high, dU_low, U0)
            return U0 + dU_high
        if self.name == "Euler":
            # Euler._step (integrators.py:396-407): U + dt f(U)
            return U0 + dt * self.rhs(U0)
        if self.name == "AB3":
            # Build-defined fixed-dt Adams-Bashforth 3, one RHS per step; the history
            # starts as (f0, f0), so step 1 is forward Euler and step 2 is AB2-consistent.
            f0 = self.rhs(U0)
            fm1, fm2 = self.hist if self.hist is not None else (f0, f0)
            self.hist = (f0, fm1)
            return U0 + dt / 12 * (23 * f0 - 16 * fm1 + 5 * fm2)
        raise NotImplementedError(self.name)


def from_mesh_tuple(points, cells, facets_int, facets_bc, tags, phys: Phys, dtype=torch.float32):
    """Build an (OracleMesh, OracleFVM) pair from the mesh generator's tuple, mirroring
    run_fvm.py:25-29 (fp32 vertices, interior facets first) for dtype=float32."""
    facets = np.concatenate([facets_int, facets_bc])
    bc_mask = np.concatenate([np.zeros(len(facets_int), bool), np.ones(len(facets_bc), bool)])
    pts = np.asarray(points)
    if dtype == torch.float32:
        pts = pts.astype(np.float32)
    mesh = OracleMesh(pts, cells, facets, bc_mask, dtype=dtype)
    dim = mesh.dim
    flags, U, dUdn = facet_bc_table(tags, dim)
    return mesh, OracleFVM(mesh, phys, flags, U, dUdn)
