"""Seeded Poisson-disk + Delaunay meshes of a channel with a circular/spherical obstacle.

The reference meshes with gmsh (``mesh_gen/mesh_2d/meshes_fvm.py:78-121``,
``mesh_gen/mesh_3d/meshes_fvm.py:8-55``), which is not available on this host.
This module is the host-side stand-in: it produces the same tuple the reference's
drivers consume (``run_fvm.py:14-34``)::

    points [V, d] float64, cells [N, d+1] int32,
    (interior_facets [F_i, d] int64, boundary_facets [F_b, d] int64), tags [F_b] str

Interior facets are the sorted vertex tuples in lexicographic order, as
``extract_interor_edges`` (``mesh_gen/mesh_2d/create_mesh.py:11-19``) emits them;
boundary facets follow, grouped by tag.  Boundary tags use the reference's names:
``"Left"`` (inlet), ``"Right"`` (farfield) and ``"NavierWall"`` (no-slip wall).

Points are sampled by grid-based parallel dart throwing (a maximal-ish Poisson-disk
set with minimum spacing ``h``), boundary points are placed at spacing ``h`` on the
outer box and on the obstacle, and a ring of ghost points outside the box keeps the
collinear/coplanar boundary points on the Delaunay hull.  Vertices and cells are then
sorted along a Morton curve so that neighbouring cells sit close in memory.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np
from scipy.spatial import Delaunay, cKDTree


class MeshGenerationError(RuntimeError):
    """Raised when the generated mesh fails a consistency check (mirrors the reference's
    ``MeshGenerationError``, ``mesh_gen/mesh_2d/create_mesh.py:7``)."""


@dataclasses.dataclass
class ChannelSpec:
    """Geometry of the channel problem: box [0,Lx]x[0,Ly](x[0,Lz]) minus a ball."""
    lengths: tuple = (2.0, 1.0)
    centre: tuple = (0.5, 0.5)
    radius: float = 0.1


# ----------------------------------------------------------------------------------
# Morton ordering
# ----------------------------------------------------------------------------------
def _part1by1(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.uint64) & np.uint64(0xFFFFFFFF)
    v = (v | (v << np.uint64(16))) & np.uint64(0x0000FFFF0000FFFF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x00FF00FF00FF00FF)
    v = (v | (v << np.uint64(4))) & np.uint64(0x0F0F0F0F0F0F0F0F)
    v = (v | (v << np.uint64(2))) & np.uint64(0x3333333333333333)
    v = (v | (v << np.uint64(1))) & np.uint64(0x5555555555555555)
    return v


def _part1by2(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.uint64) & np.uint64(0x1FFFFF)
    v = (v | (v << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
    v = (v | (v << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
    v = (v | (v << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
    return v


def morton_order(x: np.ndarray, lo: np.ndarray, hi: np.ndarray) -> np.ndarray:
    """Permutation sorting points ``x`` [n, d] along a Morton (Z-order) curve."""
    d = x.shape[1]
    bits = 21 if d == 3 else 31
    scale = (2 ** bits - 1) / np.maximum(hi - lo, 1e-300)
    q = np.clip(((x - lo) * scale), 0, 2 ** bits - 1).astype(np.uint64)
    if d == 2:
        code = _part1by1(q[:, 0]) | (_part1by1(q[:, 1]) << np.uint64(1))
    else:
        code = _part1by2(q[:, 0]) | (_part1by2(q[:, 1]) << np.uint64(1)) | (_part1by2(q[:, 2]) << np.uint64(2))
    return np.argsort(code, kind="stable")


# ----------------------------------------------------------------------------------
# Poisson-disk sampling by phased dart throwing on a background grid
# ----------------------------------------------------------------------------------
def _poisson_disk(rng: np.random.Generator, lo, hi, h: float, inside, fixed: np.ndarray,
                  rounds: int = 12) -> np.ndarray:
    """Sample points in the box [lo, hi] with pairwise distance >= h.

    ``inside(p)`` masks admissible candidates; ``fixed`` are boundary points that
    candidates must also keep a distance >= h from.  The grid has cell size h/sqrt(d)
    so every cell holds at most one sample; cells of one phase class are >= 3 cells
    apart, so candidates thrown in the same phase never conflict with each other.
    """
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    d = lo.shape[0]
    s = h / math.sqrt(d)
    shape = np.maximum(np.ceil((hi - lo) / s).astype(int), 1)
    pad = 2
    grid = np.full(tuple(shape + 2 * pad) + (d,), np.nan)
    ftree = cKDTree(fixed) if len(fixed) else None

    offs = np.stack(np.meshgrid(*[np.arange(-pad, pad + 1)] * d, indexing="ij"), -1).reshape(-1, d)
    idx_all = np.stack(np.meshgrid(*[np.arange(n) for n in shape], indexing="ij"), -1).reshape(-1, d)
    phase_of = (idx_all % 3) @ (3 ** np.arange(d))
    phases = [idx_all[phase_of == p] for p in range(3 ** d)]

    for _ in range(rounds):
        for cells in phases:
            if len(cells) == 0:
                continue
            g = tuple((cells + pad).T)
            empty = np.isnan(grid[g][:, 0])
            cells = cells[empty]
            if len(cells) == 0:
                continue
            cand = lo + (cells + rng.random((len(cells), d))) * s
            ok = inside(cand) & np.all(cand < hi, axis=1)
            for o in offs:
                nb = grid[tuple((cells + pad + o).T)]
                dist2 = np.sum((nb - cand) ** 2, axis=1)
                ok &= ~(dist2 < h * h)      # NaN neighbours compare False -> kept
            if ftree is not None and ok.any():
                dmin, _ = ftree.query(cand[ok], k=1)
                sub = np.flatnonzero(ok)
                ok[sub[dmin < h]] = False
            acc = cells[ok]
            grid[tuple((acc + pad).T)] = cand[ok]
    pts = grid.reshape(-1, d)
    return pts[~np.isnan(pts[:, 0])]


def _jittered_lattice(rng: np.random.Generator, lo, hi, h: float, inside, fixed: np.ndarray,
                      jitter: float = 0.15) -> np.ndarray:
    """Points of a cubic lattice with the Poisson-disk sampler's density (spacing
    1.16 h), each moved by a uniform jitter of +-jitter spacings, kept if admissible and
    at least 0.6 h from every fixed (boundary) point.  O(n) and vectorised: the interior
    sampler for large 3-D meshes."""
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    a = h / 0.64 ** (1.0 / 3.0)
    axes = [np.arange(l + 0.5 * a, u - 0.25 * a, a) for l, u in zip(lo, hi)]
    g = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, len(lo))
    g = g + (rng.random(g.shape) * 2.0 - 1.0) * (jitter * a)
    g = g[inside(g) & np.all(g > lo + 0.3 * h, 1) & np.all(g < hi - 0.3 * h, 1)]
    if len(fixed):
        dmin, _ = cKDTree(fixed).query(g, k=1)
        g = g[dmin >= 0.6 * h]
    return g


def _segment_points(a, b, h):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    n = max(int(round(np.linalg.norm(b - a) / h)), 1)
    t = np.arange(n + 1) / n
    return a[None] + t[:, None] * (b - a)[None]


# ----------------------------------------------------------------------------------
# 2-D channel
# ----------------------------------------------------------------------------------
def _spacing_for_cells_2d(spec: ChannelSpec, n_cells: int) -> float:
    area = spec.lengths[0] * spec.lengths[1] - math.pi * spec.radius ** 2
    # A saturated Poisson-disk set with spacing h has ~0.69/h^2 points per unit area
    # (calibrated for this sampler) and a Delaunay mesh has ~2 triangles per point.
    return math.sqrt(2.0 * 0.69 * area / n_cells) * 0.966


def channel_mesh_2d(n_cells: int, seed: int = 1, spec: ChannelSpec | None = None, h: float | None = None):
    """Triangle mesh of [0,Lx]x[0,Ly] minus a disk, with about ``n_cells`` cells."""
    spec = spec or ChannelSpec()
    Lx, Ly = spec.lengths
    cx, cy = spec.centre
    R = spec.radius
    if h is None:
        h = _spacing_for_cells_2d(spec, n_cells)
    rng = np.random.default_rng(seed)

    # Boundary points: box edges at spacing h, circle at arc spacing ~h.
    corners = [(0, 0), (Lx, 0), (Lx, Ly), (0, Ly)]
    box = np.concatenate([_segment_points(corners[i], corners[(i + 1) % 4], h)[:-1] for i in range(4)])
    nc = max(int(round(2 * math.pi * R / h)), 8)
    th = 2 * math.pi * np.arange(nc) / nc
    circ = np.stack([cx + R * np.cos(th), cy + R * np.sin(th)], 1)
    fixed = np.concatenate([box, circ])

    def inside(p):
        return ((p[:, 0] > 0) & (p[:, 0] < Lx) & (p[:, 1] > 0) & (p[:, 1] < Ly)
                & ((p[:, 0] - cx) ** 2 + (p[:, 1] - cy) ** 2 > R * R))

    interior = _poisson_disk(rng, (0, 0), (Lx, Ly), h, inside, fixed)
    pts = np.concatenate([fixed, interior])

    # Ghost ring outside the box keeps collinear box points on the hull.
    gh = np.concatenate([_segment_points((-h, -h), (Lx + h, -h), h), _segment_points((Lx + h, -h), (Lx + h, Ly + h), h),
                         _segment_points((Lx + h, Ly + h), (-h, Ly + h), h), _segment_points((-h, Ly + h), (-h, -h), h)])
    n_real = len(pts)
    tri = Delaunay(np.concatenate([pts, gh])).simplices
    tri = tri[np.all(tri < n_real, axis=1)]
    # Cells spanned by obstacle points only tile the inside of the obstacle polygon.
    on_obst = (tri >= len(box)) & (tri < len(fixed))
    tri = tri[~np.all(on_obst, axis=1)]
    area = 0.5 * np.abs((pts[tri[:, 1], 0] - pts[tri[:, 0], 0]) * (pts[tri[:, 2], 1] - pts[tri[:, 0], 1])
                        - (pts[tri[:, 1], 1] - pts[tri[:, 0], 1]) * (pts[tri[:, 2], 0] - pts[tri[:, 0], 0]))
    tri = tri[area > 1e-6 * h * h]
    return _finalise(pts, tri, spec, h, dim=2)


# ----------------------------------------------------------------------------------
# 3-D channel
# ----------------------------------------------------------------------------------
def _fibonacci_sphere(n: int) -> np.ndarray:
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    th = math.pi * (1 + 5 ** 0.5) * i
    return np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)


def channel_mesh_3d(n_cells: int, seed: int = 1, spec: ChannelSpec | None = None, h: float | None = None,
                    sampler: str = "auto"):
    """Tetrahedral mesh of [0,Lx]x[0,Ly]x[0,Lz] minus a ball, with about ``n_cells`` cells.

    sampler: "poisson" (dart throwing, slow above a few 100k cells), "lattice" (jittered
    cubic lattice inside, Poisson-disk on the faces) or "auto" (lattice above 200k)."""
    if sampler == "auto":
        sampler = "lattice" if n_cells > 200_000 else "poisson"
    if sampler not in ("poisson", "lattice"):
        raise ValueError(f"unknown sampler {sampler!r}")
    spec = spec or ChannelSpec(lengths=(2.0, 1.0, 1.0), centre=(0.5, 0.5, 0.5), radius=0.1)
    L = np.asarray(spec.lengths, float)
    c = np.asarray(spec.centre, float)
    R = spec.radius
    if h is None:
        vol = float(np.prod(L)) - 4.0 / 3.0 * math.pi * R ** 3
        # ~0.64/h^3 points per unit volume for this sampler, ~6.3 tets per point (the
        # jittered lattice gives ~10 % more tets per point).
        h = (0.64 * 6.3 * vol / n_cells) ** (1.0 / 3.0) * (1.031 if sampler == "lattice" else 1.0)
    rng = np.random.default_rng(seed)

    # Box edges, then face interiors (2-D Poisson disk per face), then the sphere.
    fixed = []
    verts = np.array([[i, j, k] for i in (0, 1) for j in (0, 1) for k in (0, 1)], float) * L
    edges = [(a, b) for a in range(8) for b in range(a + 1, 8) if np.sum(verts[a] != verts[b]) == 1]
    for a, b in edges:
        fixed.append(_segment_points(verts[a], verts[b], h))
    fixed = np.unique(np.round(np.concatenate(fixed), 12), axis=0)
    face_pts = []
    for ax in range(3):
        u, v = [a for a in range(3) if a != ax]
        for side in (0.0, L[ax]):
            sel = np.isclose(fixed[:, ax], side)
            edge2d = fixed[sel][:, [u, v]]

            def inside2(p, u=u, v=v):
                return (p[:, 0] > 0) & (p[:, 0] < L[u]) & (p[:, 1] > 0) & (p[:, 1] < L[v])

            p2 = _poisson_disk(rng, (0, 0), (L[u], L[v]), h, inside2, edge2d)
            p3 = np.zeros((len(p2), 3))
            p3[:, ax] = side
            p3[:, u] = p2[:, 0]
            p3[:, v] = p2[:, 1]
            face_pts.append(p3)
    ns = max(int(round(4 * math.pi * R * R / (0.8 * h * h))), 12)
    sph = c + R * _fibonacci_sphere(ns)
    fixed = np.concatenate([fixed] + face_pts + [sph])

    def inside(p):
        return np.all(p > 0, 1) & np.all(p < L, 1) & (np.sum((p - c) ** 2, 1) > R * R)

    if sampler == "lattice":
        interior = _jittered_lattice(rng, (0, 0, 0), L, h, inside, fixed)
    else:
        # Face points are only h apart at the least, so an interior point h from all of
        # them can still sit almost on the face and span a sliver with three of them:
        # keep interior points half a spacing off the walls and the sphere.
        def inside_off_walls(p):
            return (np.all(p > 0.5 * h, 1) & np.all(p < L - 0.5 * h, 1)
                    & (np.sum((p - c) ** 2, 1) > (R + 0.5 * h) ** 2))

        interior = _poisson_disk(rng, (0, 0, 0), L, h, inside_off_walls, fixed)
    pts = np.concatenate([fixed, interior])
    # The box is the convex hull; joggled input keeps every coplanar face point a vertex.
    tri = Delaunay(pts, qhull_options="QJ Qbb")
    if len(tri.coplanar):
        raise MeshGenerationError("Delaunay dropped coplanar points")
    tet = tri.simplices
    # Flat tets whose four vertices lie on one box face are artefacts of the hull.
    flat = np.zeros(len(tet), bool)
    for ax in range(3):
        for side in (0.0, L[ax]):
            flat |= np.all(pts[tet][:, :, ax] == side, axis=1)
    tet = tet[~flat]
    # Cells spanned by sphere points only tile the inside of the obstacle polyhedron.
    on_obst = (tet >= len(fixed) - len(sph)) & (tet < len(fixed))
    tet = tet[~np.all(on_obst, axis=1)]
    a, b, cc, dd = (pts[tet[:, i]] for i in range(4))
    vol = np.abs(np.einsum("ij,ij->i", b - a, np.cross(cc - a, dd - a))) / 6.0
    tet = tet[vol > 1e-6 * h ** 3]
    return _finalise(pts, tet, spec, h, dim=3)


# ----------------------------------------------------------------------------------
# Facet extraction, tagging and ordering
# ----------------------------------------------------------------------------------
def _face_patterns(dim):
    if dim == 2:
        return [(0, 1), (1, 2), (2, 0)]
    return [(0, 1, 2), (0, 2, 3), (0, 3, 1), (1, 3, 2)]


def _finalise(pts: np.ndarray, cells: np.ndarray, spec: ChannelSpec, h: float, dim: int):
    L = np.asarray(spec.lengths, float)
    c = np.asarray(spec.centre, float)
    R = spec.radius
    # Compact and Morton-sort vertices, then sort cells by centroid Morton code.
    used = np.unique(cells)
    remap = np.full(len(pts), -1, np.int64)
    remap[used] = np.arange(len(used))
    pts = pts[used]
    cells = remap[cells]
    lo, hi = np.zeros(dim), L
    vperm = morton_order(pts, lo, hi)
    inv = np.empty_like(vperm)
    inv[vperm] = np.arange(len(vperm))
    pts = pts[vperm]
    cells = inv[cells]
    cperm = morton_order(pts[cells].mean(axis=1), lo, hi)
    cells = cells[cperm]

    faces = np.concatenate([cells[:, list(p)] for p in _face_patterns(dim)])
    faces_sorted = np.sort(faces, axis=1)
    uniq, counts = np.unique(faces_sorted, axis=0, return_counts=True)
    if np.any(counts > 2):
        raise MeshGenerationError("non-manifold facet (shared by more than two cells)")
    interior = uniq[counts == 2]
    boundary = uniq[counts == 1]

    mid = pts[boundary].mean(axis=1)
    tol = 1e-9 * max(1.0, float(L.max()))
    tags = np.empty(len(boundary), dtype=object)
    on_left = np.abs(mid[:, 0]) < tol
    on_right = np.abs(mid[:, 0] - L[0]) < tol
    on_wall = np.zeros(len(boundary), bool)
    for ax in range(1, dim):
        on_wall |= (np.abs(mid[:, ax]) < tol) | (np.abs(mid[:, ax] - L[ax]) < tol)
    # Obstacle facets: all vertices lie on the sphere/circle.
    r_v = np.linalg.norm(pts[boundary] - c, axis=2)
    on_obst = np.all(np.abs(r_v - R) < 1e-6 * R, axis=1)
    tags[on_left] = "Left"
    tags[on_right] = "Right"
    tags[on_wall | on_obst] = "NavierWall"
    if np.any(tags == None):  # noqa: E711
        bad = int(np.sum(tags == None))  # noqa: E711
        raise MeshGenerationError(f"{bad} boundary facets are not on the domain boundary")
    # Boundary facets grouped by tag (Left, NavierWall, Right), then lexicographic.
    order_key = np.select([tags == "Left", tags == "NavierWall", tags == "Right"], [0, 1, 2])
    o = np.lexsort(tuple(boundary[:, i] for i in range(dim - 1, -1, -1)) + (order_key,))
    boundary = boundary[o]
    tags = tags[o]
    return (pts.astype(np.float64), cells.astype(np.int32),
            (interior.astype(np.int64), boundary.astype(np.int64)), [str(t) for t in tags])
