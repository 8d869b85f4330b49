"""Cell reordering for meshes that arrive in an arbitrary cell order (gmsh output, the
reference's meshes).  No reference counterpart: the reference keeps gmsh's order.

The device kernels gather neighbour records per cell (and per facet), so a wavefront of
64 consecutive cells should touch a few cache lines of neighbours, not 64 scattered
ones.  Two orders are offered:

* ``"rcm"`` -- reverse Cuthill-McKee on the cell adjacency graph (shared facets): small
  index bandwidth, the classic choice for sparse operators;
* ``"morton"`` -- cells sorted along a Z-order curve of their centroids: spatially
  compact blocks;
* ``"tiles"`` -- recursive coordinate bisection of the centroids into 256-cell leaves
  aligned with the tile-fused stage's tiles (``mesh_gen.channel.tile_order``, what the
  channel generator emits): each tile a compact box, ~30 % fewer halo cells per tile
  than a Morton order.

``reorder_cells`` permutes the cell array only (facets are vertex tuples and do not
change), so the mesh built from the result is the same mesh with renumbered cells.
States map between the two numberings with ``perm``: new cell i is old cell perm[i],
i.e. ``U_new = U_old[perm]`` and ``U_old[perm] = U_new``.
"""
from __future__ import annotations

import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import reverse_cuthill_mckee

from fvm_solver_amd.mesh_gen.channel import morton_order, tile_order


def cell_adjacency(cells: np.ndarray, n_vertices: int | None = None):
    """Symmetric cell adjacency (cells sharing a facet) as a CSR matrix."""
    cells = np.asarray(cells, np.int64)
    N, nv = cells.shape
    d = nv - 1
    if d == 2:
        pats = [(0, 1), (1, 2), (2, 0)]
    else:
        pats = [(0, 1, 2), (0, 2, 3), (0, 3, 1), (1, 3, 2)]
    faces = np.concatenate([np.sort(cells[:, list(p)], axis=1) for p in pats])
    owner = np.tile(np.arange(N), len(pats))
    order = np.lexsort(faces.T[::-1])
    fs, ow = faces[order], owner[order]
    same = np.all(fs[1:] == fs[:-1], axis=1)
    a, b = ow[:-1][same], ow[1:][same]
    A = coo_matrix((np.ones(2 * len(a)), (np.concatenate([a, b]), np.concatenate([b, a]))), shape=(N, N))
    return A.tocsr()


def bandwidth(cells: np.ndarray) -> int:
    """Largest |i - j| over adjacent cells i, j."""
    A = cell_adjacency(cells).tocoo()
    return int(np.abs(A.row - A.col).max()) if A.nnz else 0


def mean_index_distance(cells: np.ndarray) -> float:
    """Mean |i - j| over adjacent cells i, j (how far a gather reaches on average)."""
    A = cell_adjacency(cells).tocoo()
    return float(np.abs(A.row - A.col).mean()) if A.nnz else 0.0


def cell_order(points: np.ndarray, cells: np.ndarray, method: str = "rcm") -> np.ndarray:
    """Permutation ``perm`` (new -> old) of the cells."""
    cells = np.asarray(cells)
    if method == "rcm":
        return np.asarray(reverse_cuthill_mckee(cell_adjacency(cells), symmetric_mode=True), np.int64)
    if method in ("morton", "tiles"):
        pts = np.asarray(points, np.float64)
        cen = pts[cells].mean(axis=1)
        if method == "tiles":
            return tile_order(cen).astype(np.int64)
        return morton_order(cen, cen.min(0), cen.max(0)).astype(np.int64)
    raise ValueError(f"unknown cell order {method!r} (rcm, morton or tiles)")


def reorder_cells(points, cells, method: str = "rcm"):
    """(cells_new, perm): the cell array in the new order and the new -> old map."""
    perm = cell_order(points, cells, method)
    return np.asarray(cells)[perm], perm


def permuted_mesh(mesh, method: str = "tiles", device=None):
    """(mesh_new, perm): the FVMMesh with its cells renumbered (cells[perm]; vertices,
    facets and the boundary mask unchanged) and the new -> old map."""
    import torch
    pts = mesh.vertices.cpu().numpy()
    cells = mesh.cells.cpu().numpy()
    cells_new, perm = reorder_cells(pts, cells, method)
    cls = type(mesh)
    new = cls(mesh.vertices, torch.from_numpy(np.ascontiguousarray(cells_new)).to(mesh.cells.dtype), mesh.facets,
              mesh.bc_facet_mask, device=device if device is not None else mesh.device)
    return new, perm
