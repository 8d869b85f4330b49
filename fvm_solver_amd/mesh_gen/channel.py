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
collinear/coplanar boundary points on the Delaunay hull.  Large 3-D meshes use a
jittered body-centred cubic lattice instead (sliver-free Delaunay tessellation).  Vertices are
then sorted along a Morton curve and cells by a tile-aligned recursive bisection
(``tile_order``), so that neighbouring cells sit close in memory and every 256-cell tile
is a compact box.
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


def morton_codes(x: np.ndarray, lo: np.ndarray, hi: np.ndarray) -> np.ndarray:
    """Morton (Z-order) codes of points ``x`` [n, d] (d = 1, 2 or 3) in the box [lo, hi]."""
    d = x.shape[1]
    bits = 21 if d == 3 else 31
    scale = (2 ** bits - 1) / np.maximum(hi - lo, 1e-300)
    q = np.clip(((x - lo) * scale), 0, 2 ** bits - 1).astype(np.uint64)
    if d == 1:
        return q[:, 0]
    if d == 2:
        return _part1by1(q[:, 0]) | (_part1by1(q[:, 1]) << np.uint64(1))
    return _part1by2(q[:, 0]) | (_part1by2(q[:, 1]) << np.uint64(1)) | (_part1by2(q[:, 2]) << np.uint64(2))


def morton_order(x: np.ndarray, lo: np.ndarray, hi: np.ndarray) -> np.ndarray:
    """Permutation sorting points ``x`` [n, d] along a Morton (Z-order) curve."""
    return np.argsort(morton_codes(x, lo, hi), kind="stable")


def tile_order(x: np.ndarray, tile: int = 256, offset: int = 0) -> np.ndarray:
    """Permutation ordering points ``x`` [n, d] (cell centroids) by recursive coordinate
    bisection into blocks aligned with ``tile``-cell tiles: the ordered points sit at
    positions offset, offset + 1, ... of a sequence cut into tiles at multiples of
    ``tile``, and every cut of the bisection falls on a tile boundary, so each tile is
    one leaf of the bisection (a compact box of cells) and consecutive leaves are
    neighbours in space.  Against a Morton order of the same cells, a 256-cell tile has
    about 30 % fewer cells in its one- and two-ring halos (2-D and 3-D channel meshes),
    which is what the tile-fused stage stages and recomputes per tile.  Inside a tile
    the cells follow a Morton curve."""
    x = np.asarray(x, np.float64)
    n = len(x)
    if n == 0:
        return np.zeros(0, np.int64)
    code = morton_codes(x, x.min(0), x.max(0))
    out = []

    def rec(idx, a):
        b = a + len(idx)
        k_lo, k_hi = a // tile + 1, (b - 1) // tile      # tile boundaries strictly inside (a, b)
        if k_lo > k_hi:
            out.append(idx[np.argsort(code[idx], kind="stable")])
            return
        k = min(max(int(round((a + b) / (2 * tile))), k_lo), k_hi)
        nl = k * tile - a
        pts = x[idx]
        ax = int(np.argmax(pts.max(0) - pts.min(0)))
        part = np.argpartition(pts[:, ax], nl - 1, kind="introselect")
        rec(idx[part[:nl]], a)
        rec(idx[part[nl:]], k * tile)

    rec(np.arange(n, dtype=np.int64), int(offset))
    return np.concatenate(out)


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


def _bcc_lattice(rng: np.random.Generator, L: np.ndarray, a: float, jitter: float = 0.07):
    """Body-centred cubic points filling the box [0, L] with cube side a (L / a integral):
    the corner sublattice includes every face, edge and corner of the box, the centre
    sublattice sits half a cube inside.  The Delaunay tessellation of a BCC lattice is
    the disphenoid honeycomb (all tetrahedra congruent, circumspheres empty of other
    points by a margin, no four-point degeneracy), so small jitter keeps it free of
    slivers; a jittered cubic lattice, whose cubes have eight cospherical corners, is
    not.  Interior points move by +-jitter cube sides per axis; points on a box face move
    only within that face (edge points along the edge, corners not at all)."""
    n = [int(round(Lk / a)) for Lk in L]
    A = np.stack(np.meshgrid(*[np.arange(k + 1) for k in n], indexing="ij"), -1).reshape(-1, 3)
    on_face = (A == 0) | (A == np.asarray(n))
    B = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), -1).reshape(-1, 3) + 0.5
    jA = (rng.random(A.shape) * 2.0 - 1.0) * jitter
    jA[on_face] = 0.0
    jB = (rng.random(B.shape) * 2.0 - 1.0) * jitter
    return np.concatenate([(A + jA) * a, (B + jB) * a])


def _tet_quality(pts: np.ndarray, tet: np.ndarray) -> np.ndarray:
    """Inscribed-sphere-like size 3 vol / (sum of face areas) of each tetrahedron."""
    P = pts[tet]
    a, b, c, d = P[:, 0], P[:, 1], P[:, 2], P[:, 3]
    vol = np.abs(np.einsum("ij,ij->i", b - a, np.cross(c - a, d - a))) / 6.0

    def area(x, y, z):
        return 0.5 * np.linalg.norm(np.cross(y - x, z - x), axis=1)

    return 3.0 * vol / (area(a, b, c) + area(a, b, d) + area(a, c, d) + area(b, c, d))


def _circumspheres(pts: np.ndarray, tet: np.ndarray):
    """Circumcentres [n, 3] and circumradii [n] of tetrahedra."""
    P = pts[tet]
    a = P[:, 0]
    M = P[:, 1:] - a[:, None]
    rhs = 0.5 * np.sum(M * M, axis=2)
    det = np.linalg.det(M)
    ok = np.abs(det) > 1e-300
    x = np.zeros_like(a)
    x[ok] = np.linalg.solve(M[ok], rhs[ok][..., None])[..., 0]
    r = np.linalg.norm(x, axis=1)
    r[~ok] = np.inf
    return a + x, r


def _channel_tets(pts: np.ndarray, n_sph: int, L: np.ndarray):
    """Delaunay tetrahedra of pts (sphere points first) minus the flat ones of the box
    faces and those spanned by sphere points only (the obstacle's inside)."""
    tet = Delaunay(pts, qhull_options="QJ Qbb").simplices
    flat = np.zeros(len(tet), bool)
    for ax in range(3):
        for side in (0.0, L[ax]):
            flat |= np.all(np.abs(pts[tet][:, :, ax] - side) < 1e-12, axis=1)
    tet = tet[~flat]
    return tet[~np.all(tet < n_sph, axis=1)]


def _bcc_channel(rng: np.random.Generator, L: np.ndarray, c: np.ndarray, R: float, n_cells: int,
                 min_quality: float = 0.15, rounds: int = 8):
    """Points and tetrahedra of the box minus a ball: a BCC lattice with ~n_cells
    Delaunay tetrahedra, lattice points within 0.4 spacings of the ball removed, the ball's
    surface sampled by a Fibonacci spiral.  Slivers can only form where the spiral meets
    the lattice: the lattice vertices of every tetrahedron whose size 3 vol / area is
    below min_quality of the mean are moved by a random quarter spacing and the ball's
    neighbourhood B is triangulated again, until none is left.  Moving points near the
    ball changes the Delaunay tessellation only inside B (tetrahedra whose circumsphere
    contains an old or new position), so the global tessellation is computed once: its
    tetrahedra whose circumsphere leaves B are kept, the rest come from the tessellation
    of the points in B."""
    vol = float(np.prod(L)) - 4.0 / 3.0 * math.pi * R ** 3
    # the disphenoid honeycomb has 12 tetrahedra per cube of side a; the box lengths must
    # be whole numbers of cubes
    m = max(2, int(round(float(np.min(L)) / (12.0 * vol / n_cells) ** (1.0 / 3.0))))
    a = float(np.min(L)) / m
    if not np.allclose(L / a, np.round(L / a)):
        raise ValueError("the BCC sampler needs box lengths that are whole multiples of the shortest one")
    h = a * math.sqrt(3.0) / 2.0           # nearest-neighbour distance of the lattice
    P = _bcc_lattice(rng, L, a)
    P = P[np.sum((P - c) ** 2, 1) > (R + 0.4 * h) ** 2]
    ns = max(int(round(4 * math.pi * R * R / (0.9 * h * h))), 12)
    sph = c + R * _fibonacci_sphere(ns)
    pts = np.concatenate([sph, P])
    tet = _channel_tets(pts, ns, L)
    q = _tet_quality(pts, tet)
    qmean = float(q.mean())
    RB = R + 8.0 * h                      # the neighbourhood B of the ball
    cc, r = _circumspheres(pts, tet)
    inside_B = np.linalg.norm(cc - c, axis=1) + r < RB
    if np.any(q[~inside_B] < min_quality * qmean):
        raise MeshGenerationError("sliver away from the obstacle (lattice jitter too large?)")
    keep_far = tet[~inside_B]
    # local points: everything inside B (the tessellation of B's interior depends on them
    # only); lattice points beyond R + 3 spacings are never moved
    loc = np.flatnonzero(np.sum((pts - c) ** 2, 1) < RB * RB)
    loc = loc[np.argsort(loc >= ns, kind="stable")]        # sphere points first
    lp = pts[loc]
    for _ in range(rounds):
        lt = _channel_tets(lp, ns, L)
        lcc, lr = _circumspheres(lp, lt)
        lt = lt[np.linalg.norm(lcc - c, axis=1) + lr < RB]
        bad = _tet_quality(lp, lt) < min_quality * qmean
        if not bad.any():
            break
        # move the lattice vertices of every sliver by a random quarter spacing (a
        # sliver is four nearly cospherical points: a perturbation breaks it), keeping
        # them off the ball and inside B
        mv = np.unique(lt[bad][lt[bad] >= ns])
        if len(mv) == 0:
            raise MeshGenerationError("a sliver at the obstacle has no lattice vertex")
        for _try in range(20):
            step = rng.standard_normal((len(mv), 3))
            new = lp[mv] + 0.25 * h * step / np.linalg.norm(step, axis=1, keepdims=True)
            dist = np.linalg.norm(new - c, axis=1)
            ok = (dist > R + 0.4 * h) & (dist < R + 3.0 * h) & np.all(new > 0, 1) & np.all(new < L, 1)
            lp[mv[ok]] = new[ok]
            mv = mv[~ok]
            if len(mv) == 0:
                break
    else:
        raise MeshGenerationError("slivers at the obstacle remain")
    pts = pts.copy()
    pts[loc] = lp
    return pts, np.concatenate([keep_far, loc[lt]]), h


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

    sampler: "poisson" (dart throwing inside, Poisson-disk on the faces; slow above a few
    100k cells), "bcc" (jittered body-centred cubic lattice aligned with the box, slivers
    at the ball removed; ``h`` is ignored) or "auto" (bcc above 200k cells)."""
    if sampler == "auto":
        sampler = "bcc" if n_cells > 200_000 else "poisson"
    if sampler not in ("poisson", "bcc"):
        raise ValueError(f"unknown sampler {sampler!r}")
    spec = spec or ChannelSpec(lengths=(2.0, 1.0, 1.0), centre=(0.5, 0.5, 0.5), radius=0.1)
    L = np.asarray(spec.lengths, float)
    c = np.asarray(spec.centre, float)
    R = spec.radius
    rng = np.random.default_rng(seed)
    if sampler == "bcc":
        pts, tet, h = _bcc_channel(rng, L, c, R, n_cells)
        return _finalise(pts, tet, spec, h, dim=3)
    if h is None:
        vol = float(np.prod(L)) - 4.0 / 3.0 * math.pi * R ** 3
        # ~0.64/h^3 points per unit volume for this sampler, ~6.3 tets per point
        h = (0.64 * 6.3 * vol / n_cells) ** (1.0 / 3.0)

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

    # Face points are only h apart at the least, so an interior point h from all of them
    # can still sit almost on the face and span a sliver with three of them: keep
    # interior points half a spacing off the walls and the sphere.
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
    # Compact and Morton-sort vertices, then order cells in tile-aligned blocks.
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
    cperm = tile_order(pts[cells].mean(axis=1))
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
