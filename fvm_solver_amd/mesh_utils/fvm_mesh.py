"""Mesh preprocessing -- drop-in for ``time_fvm/mesh_utils/fvm_mesh.py``.

``FVMMesh2D`` / ``FVMMesh3D`` take the reference's constructor arguments
(fvm_mesh.py:115, 350) and expose the same attributes, with the same index orderings:

* ``cell_to_facet`` in local face-pattern order ((v0v1),(v1v2),(v2v0) in 2-D;
  (012),(023),(031),(132) in 3-D), ``cell_facet_signs`` = sign(n_hat . (m_f - c)),
* ``facet_to_cell_main`` = [s=+1 cell, s=-1 cell] per interior facet in facet order,
  ``facet_to_cell_bc`` per boundary facet,
* ``cell_grad_stuff`` = (cell_disps, facet_dist_bc, G_mats, combined_neigh) with the
  weighted least-squares gradient G = A^-1 d^T W^2, w = |d|^-1/4 (x7.5 on boundary
  offsets) and the diagonal -sum_k G accumulated in fp32 in local order.

Differences from the reference, by design: the work is vectorised (no per-cell Python
loops, so 1M-4M cells take seconds), and every geometric quantity -- normals, volumes,
centroids, displacements, the LSQ system and its inverse -- is computed in float64
from the given vertices and rounded to float32 only where the reference stores float32.
For a mesh on a GPU device the float64 arithmetic runs in the library's kernels
(csrc/geom.hip); the integer topology (facet lookup, left/right ordering) stays in
numpy.  The float64 arrays are kept in ``geo64`` for the device-layout builder.
Quirk kept: the 3-D normal is the full cross product, |n| = 2 x facet area
(fvm_mesh.py:418).
"""
from __future__ import annotations

import numpy as np
import torch

FACE_PATTERNS = {2: ((0, 1), (1, 2), (2, 0)), 3: ((0, 1, 2), (0, 2, 3), (0, 3, 1), (1, 3, 2))}


def _np(a, dtype=None):
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    a = np.asarray(a)
    return a.astype(dtype) if dtype is not None else a


def _sorted_keys(f: np.ndarray, nv: int) -> np.ndarray:
    f = np.sort(f.astype(np.int64), axis=1)
    key = f[:, 0].copy()
    for j in range(1, f.shape[1]):
        key = key * nv + f[:, j]
    return key


class FVMMesh:
    dim: int

    def __init__(self, vertices, cells, facets, bc_facet_mask, device="cuda"):
        self.device = device
        self.vertices = vertices if isinstance(vertices, torch.Tensor) else torch.as_tensor(vertices)
        self.cells = cells if isinstance(cells, torch.Tensor) else torch.as_tensor(cells)
        self.facets = facets if isinstance(facets, torch.Tensor) else torch.as_tensor(facets)
        self.bc_facet_mask = (bc_facet_mask if isinstance(bc_facet_mask, torch.Tensor)
                              else torch.as_tensor(bc_facet_mask)).bool()
        V = _np(vertices, np.float64)
        C = _np(cells, np.int64)
        Fv = _np(facets, np.int64)
        bcm = _np(bc_facet_mask).astype(bool)
        if Fv.shape[0] != bcm.shape[0]:
            raise ValueError(f"Different number of facets from bc facet mask {Fv.shape} {bcm.shape}")
        if C.shape[1] != self.dim + 1 or Fv.shape[1] != self.dim:
            raise ValueError("cells/facets do not match the mesh dimension")
        self.n_cells = N = C.shape[0]
        self.n_facets = F = Fv.shape[0]
        self.n_bc_facet = int(bcm.sum())
        d = self.dim
        nl = d + 1

        # ---- geometry (fp64; on the GPU when the mesh lives there) ----
        gpu = _GpuGeometry(V, d, device) if _on_gpu(device) else None
        if gpu is not None:
            n64, nh64, mid64 = gpu.facets(Fv)
            vol64, cen64 = gpu.cells(C)
        else:
            fv = V[Fv]
            if d == 2:
                e = fv[:, 1] - fv[:, 0]
                n64 = np.stack([e[:, 1], -e[:, 0]], 1)
            else:
                n64 = np.cross(fv[:, 1] - fv[:, 0], fv[:, 2] - fv[:, 0])
            mid64 = fv.mean(axis=1)
            cp = V[C]
            if d == 2:
                ab, ac = cp[:, 1] - cp[:, 0], cp[:, 2] - cp[:, 0]
                vol64 = 0.5 * np.abs(ab[:, 0] * ac[:, 1] - ab[:, 1] * ac[:, 0])
            else:
                a = cp[:, 0]
                vol64 = np.abs(np.einsum("ij,ij->i", cp[:, 1] - a, np.cross(cp[:, 2] - a, cp[:, 3] - a))) / 6.0
            cen64 = cp.mean(axis=1)
            nh64 = n64 / np.linalg.norm(n64, axis=1, keepdims=True)
        if np.any(vol64 <= 0):
            raise ValueError("degenerate cell (zero volume)")

        # ---- cell -> facet lookup (fvm_mesh.py:185-225, 435-473) ----
        nv = int(V.shape[0])
        fkey = _sorted_keys(Fv, nv)
        order = np.argsort(fkey, kind="stable")
        sk = fkey[order]
        if np.any(sk[1:] == sk[:-1]):
            raise ValueError("duplicate facets in the facet list")
        ck = np.stack([_sorted_keys(C[:, list(p)], nv) for p in FACE_PATTERNS[d]], 1)
        pos = np.minimum(np.searchsorted(sk, ck), F - 1)
        if np.any(sk[pos] != ck):
            raise KeyError("a cell facet is missing from the facet list")
        c2f = order[pos]

        # ---- orientation signs and left/right ordering (fvm_mesh.py:74-109) ----
        if gpu is not None:
            disp64, signs = gpu.disp_signs(c2f)
        else:
            disp64 = mid64[c2f] - cen64[:, None, :]
            signs = np.sign(np.einsum("nkd,nkd->nk", nh64[c2f], disp64)).astype(np.int64)
        flat_f = c2f.reshape(-1)
        counts = np.bincount(flat_f, minlength=F)
        if np.any(counts == 0) or np.any(counts > 2):
            raise ValueError("every facet must belong to one or two cells")
        if np.any((counts == 1) != bcm):
            raise ValueError("Inconsistent boundary bug: boundary mask does not match the facet/cell incidence")
        o = np.argsort(flat_f, kind="stable")
        start = np.concatenate([[0], np.cumsum(counts)[:-1]])
        two = counts == 2
        fa = o[start]
        fb = np.where(two, o[np.minimum(start + 1, len(o) - 1)], -1)
        ca, cb = fa // nl, np.where(two, fb // nl, -1)
        sa = signs.reshape(-1)[fa]
        sb = np.where(two, signs.reshape(-1)[np.maximum(fb, 0)], 0)
        flip = two & (sa == -1) & (sb == 1)
        left = np.where(flip, cb, ca)
        right = np.where(flip, ca, cb)

        # ---- least-squares gradient weights (fvm_mesh.py:263-344, 507-582) ----
        bc_local = np.full(F, -1, np.int64)
        bc_local[bcm] = np.arange(self.n_bc_facet)
        lf, rf = left[c2f], right[c2f]
        other = np.where(lf == np.arange(N)[:, None], rf, lf)
        is_bc = bcm[c2f]
        combined = np.where(is_bc, bc_local[c2f] + N, other)
        if gpu is not None:
            G64 = gpu.lsq(c2f, other, is_bc)
        else:
            npts = np.where(is_bc[..., None], mid64[c2f], cen64[np.maximum(other, 0)])
            dv = npts - cen64[:, None, :]
            w = np.linalg.norm(dv, axis=2) ** -0.25
            w[is_bc] *= 7.5
            w2 = w * w
            A = np.einsum("nk,nki,nkj->nij", w2, dv, dv)
            G64 = np.einsum("nij,nkj,nk->nik", np.linalg.inv(A), dv, w2)      # [N, d, nl]

        self.geo64 = dict(normals=n64, normals_hat=nh64, midpoints=mid64, volumes=vol64, centroids=cen64,
                          cent_to_facet_disp=disp64, G=G64, left=left, right=right, bc_local=bc_local,
                          is_bc=is_bc, c2f=c2f, signs=signs, combined=combined)

        # ---- reference-named attributes ----
        f32 = torch.float32
        self.normals = torch.from_numpy(n64).to(f32)
        self.midpoints = torch.from_numpy(mid64).to(f32)
        self.volumes = torch.from_numpy(vol64).to(f32)
        self.centroids = torch.from_numpy(cen64).to(f32)
        self.cell_to_facet = torch.from_numpy(c2f)
        self.cell_facet_signs = torch.from_numpy(signs)
        self.cent_to_facet_disp = torch.from_numpy(disp64).to(f32)
        self.facet_to_cell_main = torch.from_numpy(np.stack([left[~bcm], right[~bcm]], 1))
        self.facet_to_cell_bc = torch.from_numpy(left[bcm])
        self.normals_main = self.normals[torch.from_numpy(~bcm)]
        self._grad_stuff = None

    @property
    def cell_grad_stuff(self):
        """(cell_disps [F_i, d], facet_dist_bc [F_b], G_mats: d COO [N, N+F_b], combined_neigh [N, nl])."""
        if self._grad_stuff is None:
            g = self.geo64
            bcm = self.bc_facet_mask.numpy()
            N, d, nl = self.n_cells, self.dim, self.dim + 1
            cen = g["centroids"]
            cell_disps = torch.from_numpy(cen[g["right"][~bcm]] - cen[g["left"][~bcm]]).float()
            fp = g["midpoints"][bcm] - cen[g["left"][bcm]]
            dist = torch.from_numpy(np.einsum("nd,nd->n", fp, g["normals_hat"][bcm])).float()
            G32 = torch.from_numpy(g["G"]).float()
            comb = torch.from_numpy(g["combined"]).int()
            rows = torch.arange(N)
            G_mats = []
            for i in range(d):
                diag = G32[:, i, 0].clone()
                for k in range(1, nl):
                    diag = diag + G32[:, i, k]
                r = torch.cat([rows[:, None].expand(N, nl), rows[:, None]], 1).reshape(-1)
                c = torch.cat([comb.long(), rows[:, None]], 1).reshape(-1)
                v = torch.cat([G32[:, i, :], -diag[:, None]], 1).reshape(-1)
                G_mats.append(torch.sparse_coo_tensor(torch.stack([r, c]), v, (N, N + self.n_bc_facet)))
            self._grad_stuff = (cell_disps, dist, G_mats, comb)
        return self._grad_stuff


def _on_gpu(device) -> bool:
    try:
        return torch.device(device).type == "cuda" and torch.cuda.is_available()
    except (RuntimeError, TypeError):
        return False


class _GpuGeometry:
    """The mesh's fp64 geometry computed by the library's kernels (csrc/geom.hip); the
    results come back as numpy arrays for the host-side layout builder."""

    def __init__(self, V: np.ndarray, dim: int, device):
        from fvm_solver_amd import _native as nat
        self.nat, self.dim, self.dev = nat, dim, torch.device(device)
        self.V = torch.from_numpy(np.ascontiguousarray(V, np.float64)).to(self.dev)

    def _t(self, a, dt=torch.int64):
        return torch.from_numpy(np.ascontiguousarray(a)).to(self.dev, dt)

    def facets(self, Fv):
        nat, F, d = self.nat, Fv.shape[0], self.dim
        fv = self._t(Fv)
        n, nh, mid = (torch.empty(F, d, dtype=torch.float64, device=self.dev) for _ in range(3))
        nat.check(nat.lib().fvm_geom_facets_f64(nat.ptr(self.V), d, nat.ptr(fv), F, nat.ptr(n), nat.ptr(nh),
                                                nat.ptr(mid), nat.stream_ptr(self.dev)), "fvm_geom_facets_f64")
        self.mid, self.nh = mid, nh
        return n.cpu().numpy(), nh.cpu().numpy(), mid.cpu().numpy()

    def cells(self, Cv):
        nat, N, d = self.nat, Cv.shape[0], self.dim
        cv = self._t(Cv)
        vol = torch.empty(N, dtype=torch.float64, device=self.dev)
        cen = torch.empty(N, d, dtype=torch.float64, device=self.dev)
        nat.check(nat.lib().fvm_geom_cells_f64(nat.ptr(self.V), d, nat.ptr(cv), N, nat.ptr(vol), nat.ptr(cen),
                                               nat.stream_ptr(self.dev)), "fvm_geom_cells_f64")
        self.cen = cen
        return vol.cpu().numpy(), cen.cpu().numpy()

    def disp_signs(self, c2f):
        nat, (N, nl), d = self.nat, c2f.shape, self.dim
        self.c2f = self._t(c2f)
        disp = torch.empty(N, nl, d, dtype=torch.float64, device=self.dev)
        sg = torch.empty(N, nl, dtype=torch.int64, device=self.dev)
        nat.check(nat.lib().fvm_geom_disp_signs_f64(nat.ptr(self.cen), nat.ptr(self.mid), nat.ptr(self.nh), d,
                                                    nat.ptr(self.c2f), N, nat.ptr(disp), nat.ptr(sg),
                                                    nat.stream_ptr(self.dev)), "fvm_geom_disp_signs_f64")
        return disp.cpu().numpy(), sg.cpu().numpy()

    def lsq(self, c2f, other, is_bc):
        nat, (N, nl), d = self.nat, c2f.shape, self.dim
        G = torch.empty(N, d, nl, dtype=torch.float64, device=self.dev)
        ot, bc = self._t(other), self._t(is_bc.astype(np.uint8), torch.uint8)
        nat.check(nat.lib().fvm_geom_lsq_f64(nat.ptr(self.cen), nat.ptr(self.mid), d, nat.ptr(self.c2f),
                                             nat.ptr(ot), nat.ptr(bc), N, nat.ptr(G), nat.stream_ptr(self.dev)),
                  "fvm_geom_lsq_f64")
        return G.cpu().numpy()


class FVMMesh2D(FVMMesh):
    """Triangle mesh (fvm_mesh.py:112-344)."""
    dim = 2


class FVMMesh3D(FVMMesh):
    """Tetrahedral mesh (fvm_mesh.py:347-582)."""
    dim = 3


def mesh_from_tuple(points, cells, facets_int, facets_bc, device="cuda"):
    """Assemble an FVMMesh from the mesh generator's tuple the way the reference drivers
    do (run_fvm.py:25-29): float32 vertices, interior facets first, then boundary."""
    facets = np.concatenate([np.asarray(facets_int), np.asarray(facets_bc)]).astype(np.int64)
    bc = np.concatenate([np.zeros(len(facets_int), bool), np.ones(len(facets_bc), bool)])
    Xs = torch.from_numpy(np.asarray(points)).float()
    cls = FVMMesh2D if np.asarray(cells).shape[1] == 3 else FVMMesh3D
    return cls(Xs, torch.from_numpy(np.asarray(cells)).int(), torch.from_numpy(facets), torch.from_numpy(bc),
               device=device)
