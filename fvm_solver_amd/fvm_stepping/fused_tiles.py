"""Tile tables of the tile-fused RK stage kernel (``k_fused`` in csrc/rhs.hip).

The fused kernel gives each workgroup one tile of 256 consecutive cells and evaluates,
in LDS, everything a stage needs for those cells: the gradient and limiter of the tile's
cells and of their one-ring halo (H1, the cells outside the tile that share a facet with
it), the fluxes of every facet of the tile's cells, and the divergence + stage update.
H1 gradients need the primitives of H1's own neighbours (H2), so the tile stages
primitives of tile + H1 + H2.  No reference counterpart: the reference evaluates the
same operators as separate SpMV/SpMM and elementwise passes over the whole mesh
(facet_process.py:332-450, fvm_equation.py:406-435).

Local cell numbering of tile t (cells c0 = 256 t ...):
    0 .. 255           the tile's own cells (c0 + l; a short last tile leaves a gap)
    256 .. 256+n1-1    H1, ascending global id
    256+n1 .. +n2-1    H2, ascending global id

Facets of tile t: its *owned* facets -- those whose lowest adjacent cell is in the tile,
a contiguous run [fo0, fo0 + nfo) of the device facet order (grouped by the tile of the
lowest adjacent cell) -- then its *foreign* facets, interior facets of a tile cell owned by an earlier
tile (evaluated by both tiles, each keeping its own cell's share).

Facet descriptor (uint32): bits 0-11 local id of the slot-0 cell, 12-13 that cell's slot
index k (cell_facet[k] == f), 14-25 / 26-27 the same for the slot-1 cell, bit 28 the
slot-0 cell has the lower global id (A_face_grad's ascending column order), bit 29
boundary facet, bit 30 the outer (boundary) side is slot 0.  For a boundary facet the
outer side's local id is 0xFFF.

Face values live in LDS per facet of the list and side, FF[q][side][C]: the gradient
phase writes them from the cell on that side (a tile cell finds q through its cfac row
and its side through its slot bits; an H1 cell through the H1 facet table h1f), the
facet's flux then overwrites FF[q][0].  Tiles hold 256 cells in 2-D and 3-D; a thread
evaluates up to 2 (2-D) or 3 (3-D) facets of its tile's list.
"""
from __future__ import annotations

import numpy as np

TILE = 256
TAB = 12          # ints per tile in the tile table
NONE16 = 0xFFFF   # local neighbour id of a boundary slot; H1 slot whose facet is not in the list
LOCAL_MAX = 0xFFF
SIDE_BIT = 0x8000  # H1 facet table: the H1 cell is the facet's slot-1 side
TILE_MESHES = 4    # meshes of a batch whose physics a tile keeps in LDS


def facets_per_thread(D: int) -> int:
    """Facets each of the 256 threads of a tile evaluates at most (2-D: 2, 3-D: 3)."""
    return 2 if D == 2 else 3


def fused_lds_bytes(D: int, ncl: int, ngl: int, nfl: int, batch: bool) -> int:
    """Dynamic LDS of the fused kernel (the same formula as the kernel's launch): primitives
    Pl [ncl][C], limited V/T gradients GL [D][D+1][ngl], both sides' face values per
    facet FF [nfl][2][C] (the facet's flux overwrites its side-0 values), each block
    starting on a 16-byte boundary; for a batch the tile cells' 16-bit mesh ids and a
    4 x 4-float mesh table."""
    C = D + 2
    r4 = lambda x: (x + 3) & ~3
    floats = r4(ncl * C) + r4(ngl * D * (D + 1)) + nfl * 2 * C
    return 4 * floats + ((TILE * 2 + TILE_MESHES * 4 * 4) if batch else 0)


def tangent_basis(nh) -> np.ndarray:
    """Unit tangent vectors [F][D-1][D] of the unit normals nh [F][D] (fp64): in 2-D
    (-n_y, n_x); in 3-D the basis of Duff et al., "Building an Orthonormal Basis,
    Revisited" (JCGT 6(1), 2017), with sign(+-0) = +1 -- the basis the kernels rebuild
    from their fp32 n_hat."""
    nh = np.asarray(nh, np.float64)
    if nh.shape[1] == 2:
        return np.stack([-nh[:, 1], nh[:, 0]], 1)[:, None, :]
    sg = np.where(nh[:, 2] >= 0, 1.0, -1.0)
    a = -1.0 / (sg + nh[:, 2])
    b = nh[:, 0] * nh[:, 1] * a
    t1 = np.stack([1.0 + sg * nh[:, 0] ** 2 * a, sg * b, -sg * nh[:, 0]], 1)
    t2 = np.stack([b, sg + nh[:, 1] ** 2 * a, -nh[:, 1]], 1)
    return np.stack([t1, t2], 1)


def tangent_coords(geo, D: int) -> np.ndarray:
    """X_orthog's coordinates [F][D-1] in the tangent basis of n_hat, from the full facet
    record (n_hat at columns D..2D-1, X_orthog at 2D+2..3D+1).  X_orthog = d / (n_hat.d) -
    n_hat is orthogonal to n_hat, so these coordinates carry all of it."""
    geo = np.asarray(geo, np.float64)
    return np.einsum("fjd,fd->fj", tangent_basis(geo[:, D:2 * D]), geo[:, 2 * D + 2:3 * D + 2])


def compact_geometry(geo, D: int) -> np.ndarray:
    """The fused stage's facet record from the full one (n [D], n_hat [D], |n|, 1/(n_hat.d),
    X_orthog [D], ...): n, 1/(n_hat.d), X_orthog's tangent coordinates -- 2 D floats."""
    geo = np.asarray(geo, np.float64)
    out = np.concatenate([geo[:, :D], geo[:, 2 * D + 1:2 * D + 2], tangent_coords(geo, D)], 1)
    return np.ascontiguousarray(out.astype(np.float32))


def device_geometry(geo, D: int) -> np.ndarray:
    """The full facet record as the kernels read it: X_orthog's D columns hold its D - 1
    tangent coordinates, then 0 (the kernels rebuild X_orthog from them)."""
    g = np.array(geo, np.float64, copy=True)
    g[:, 2 * D + 2:3 * D + 2] = 0.0
    g[:, 2 * D + 2:3 * D + 1] = tangent_coords(geo, D)
    return g


NONE12 = 0xFFF     # packed slot word: boundary slot (no local neighbour)
NONE10 = 0x3FF     # packed slot word: facet not in the tile's list (H1 cells)


def hrec_n(D: int) -> int:
    """Floats per H1 record: LSQ values of the NL neighbour columns, the displacements of
    the first NL - 1 slots (the last is minus their sum), eps."""
    return (D + 1) * D + D * D + 1


def pack_slot_words(nl, q, side, ref):
    """One 32-bit word per cell slot for the fused kernel: bits 0-11 the local id of the
    neighbour behind the slot (NONE12 for a boundary facet), 12-21 the facet's position
    in the tile's list (NONE10: not in it), 22 the cell is the facet's slot-1 side, 23-24
    the slot's position in the reference's local order of the cell's LSQ row."""
    nl, q, side, ref = (np.asarray(a, np.int64) for a in (nl, q, side, ref))
    if np.any((nl < 0) | (nl > NONE12)) or np.any((q < 0) | (q > NONE10)) or np.any((ref < 0) | (ref > 3)):
        raise TilesUnavailable("slot word field out of range")
    return (nl | (q << 12) | (side << 22) | (ref << 23)).astype(np.uint32)


def unpack_slot_words(w):
    """Inverse of pack_slot_words: (nl, q, side, ref)."""
    w = np.asarray(w, np.int64)
    return w & 0xFFF, (w >> 12) & 0x3FF, (w >> 22) & 1, (w >> 23) & 3


class TilesUnavailable(ValueError):
    pass


def _lookup(keys: np.ndarray, q: np.ndarray) -> np.ndarray:
    """Positions of q in the sorted unique array keys; -1 where absent."""
    pos = np.searchsorted(keys, q)
    pos = np.minimum(pos, max(len(keys) - 1, 0))
    ok = (len(keys) > 0) & (keys[pos] == q) if len(keys) else np.zeros(len(q), bool)
    return np.where(ok, pos, -1)


def build_tiles(N: int, D: int, nbr: np.ndarray, cell_facet: np.ndarray, fac_cell: np.ndarray,
                gval: np.ndarray, disp: np.ndarray, eps: np.ndarray, geo: np.ndarray,
                bcell: np.ndarray, group: np.ndarray, cell_mesh: np.ndarray, group_off: np.ndarray,
                lds_limit: int = 64 * 1024, lds_target: int | None = None, gref: np.ndarray | None = None) -> dict:
    """Tile tables for the fused kernel.

    nbr [N, NL] neighbour per slot (>= N: N + boundary index); cell_facet [N, NL] device
    facet positions; fac_cell [2, F] device order; gval [N, D, NL+1]; disp [N, NL, D];
    eps [N]; geo [F, G] device order (first 3D+2 floats: n, n_hat, |n|, 1/(n.d),
    X); bcell/group [Fb]; cell_mesh [N]; group_off = first group of each mesh (+ total);
    gref [N, NL] the position of each slot's column in the reference's local order of
    the cell's G_mats row (the order its own column, minus the sum of the others, is
    summed in; default: slot order).
    Raises TilesUnavailable when a tile's working set does not fit ``lds_limit`` bytes
    of LDS or the local ids overflow (a badly ordered mesh).  ``lds_target`` (optional,
    <= lds_limit) is only reported: whether the tables fit it (e.g. the LDS that lets
    one more workgroup share a CU)."""
    NL = D + 1
    C = D + 2
    nbr = np.asarray(nbr, np.int64)
    cell_facet = np.asarray(cell_facet, np.int64)
    fac_cell = np.asarray(fac_cell, np.int64)
    F = fac_cell.shape[1]
    T = (N + TILE - 1) // TILE
    tile_of = np.arange(N, dtype=np.int64) // TILE

    # ---- H1: (tile, cell) pairs across tile borders ----
    src = np.repeat(np.arange(N, dtype=np.int64), NL)
    dst = nbr.reshape(-1)
    inner = dst < N
    a, b = src[inner], dst[inner]
    cross = tile_of[a] != tile_of[b]
    K1 = np.unique(tile_of[a[cross]] * N + b[cross])
    t1 = K1 // N
    n1 = np.bincount(t1, minlength=T).astype(np.int64)
    s1 = np.concatenate([[0], np.cumsum(n1)])
    # ---- H2: neighbours of H1 outside tile and H1 ----
    h1c = K1 % N
    cand_t = np.repeat(t1, NL)
    cand_c = nbr[h1c].reshape(-1)
    ok = cand_c < N
    cand_t, cand_c = cand_t[ok], cand_c[ok]
    ok = tile_of[cand_c] != cand_t
    ck = np.unique(cand_t[ok] * N + cand_c[ok])
    K2 = ck[_lookup(K1, ck) < 0]
    t2 = K2 // N
    n2 = np.bincount(t2, minlength=T).astype(np.int64)
    s2 = np.concatenate([[0], np.cumsum(n2)])

    if T and int((n1 + n2).max(initial=0)) + TILE > LOCAL_MAX:
        raise TilesUnavailable("tile halo too large for 12-bit local ids (cells not ordered along a curve?)")

    def local(t, c):
        """Local ids of global cells c in tiles t (vectorised); -1 if not staged."""
        t = np.asarray(t, np.int64)
        c = np.asarray(c, np.int64)
        out = np.full(c.shape, -1, np.int64)
        own = tile_of[np.minimum(c, N - 1)] == t
        own &= c < N
        out[own] = c[own] - t[own] * TILE
        rest = ~own & (c < N)
        key = t[rest] * N + c[rest]
        p1 = _lookup(K1, key)
        p2 = _lookup(K2, key)
        tr = t[rest]
        loc = np.where(p1 >= 0, TILE + p1 - s1[tr], np.where(p2 >= 0, TILE + n1[tr] + p2 - s2[tr], -1))
        out[rest] = loc
        return out

    # ---- local neighbour rows of own + H1 cells ----
    nrow = TILE + n1
    nbr_off = np.concatenate([[0], np.cumsum(nrow)])
    rows = np.full((int(nbr_off[-1]), 4), NONE16, np.uint16)
    # own cells (rows of a short last tile beyond N stay NONE16)
    own_row = nbr_off[tile_of] + (np.arange(N) - tile_of * TILE)
    ln = local(np.repeat(tile_of, NL), nbr.reshape(-1)).reshape(N, NL)
    if ((ln < 0) & (nbr < N)).any():
        raise AssertionError("own neighbour not staged")
    rows[own_row, :NL] = np.where(nbr < N, ln, NONE16).astype(np.uint16)
    # H1 cells
    h1_row = nbr_off[t1] + TILE + (np.arange(len(K1)) - s1[t1])
    lh = local(np.repeat(t1, NL), nbr[h1c].reshape(-1)).reshape(-1, NL)
    if ((lh < 0) & (nbr[h1c] < N)).any():
        raise AssertionError("H1 neighbour not staged")
    rows[h1_row, :NL] = np.where(nbr[h1c] < N, lh, NONE16).astype(np.uint16)

    # ---- H1 static records: LSQ values of the neighbour columns [NL][D] (slot order; the
    # kernel derives the cell's own column as the reference does), displacements of the
    # first NL - 1 slots [NL-1][D] (the last is minus their sum), eps, stored
    # component-major [HN][sum n1] (the kernel's lanes read one row per component) ----
    HN = hrec_n(D)
    rec = np.zeros((len(K1), HN), np.float32)
    rec[:, :NL * D] = np.asarray(gval, np.float32)[h1c][:, :, :NL].transpose(0, 2, 1).reshape(len(K1), -1)
    o = NL * D
    rec[:, o:o + (NL - 1) * D] = np.asarray(disp, np.float32)[h1c][:, :NL - 1].reshape(len(K1), -1)
    rec[:, o + (NL - 1) * D] = np.asarray(eps, np.float32)[h1c]
    hrec = np.ascontiguousarray(rec.T)

    # ---- facets ----
    lo = fac_cell.min(0)
    hi = fac_cell.max(0)
    if np.any(np.diff(lo // TILE) < 0):
        raise ValueError("facets are not in device order (ascending tile of the lowest cell)")
    own_t = lo // TILE
    fo0 = np.searchsorted(own_t, np.arange(T), side="left")
    fo1 = np.searchsorted(own_t, np.arange(T), side="right")
    nfo = fo1 - fo0
    interior = hi < N
    fx = np.flatnonzero(interior & (tile_of[np.minimum(hi, N - 1)] != own_t))
    ft = tile_of[hi[fx]]
    order = np.lexsort((fx, ft))
    fx, ft = fx[order], ft[order]
    nfor = np.bincount(ft, minlength=T).astype(np.int64)
    foff = np.concatenate([[0], np.cumsum(nfor)])

    # slot index of each facet side
    side_slot = np.full((2, F), 0, np.int64)
    cidx = np.repeat(np.arange(N, dtype=np.int64), NL)
    kidx = np.tile(np.arange(NL, dtype=np.int64), N)
    fl = cell_facet.reshape(-1)
    sv = np.where(fac_cell[0, fl] == cidx, 0, 1)
    if np.any(fac_cell[sv, fl] != cidx):
        raise AssertionError("cell_facet and fac_cell disagree")
    side_slot[sv, fl] = kidx

    def descriptors(t, f):
        c0_, c1_ = fac_cell[0, f], fac_cell[1, f]
        bnd = (c0_ >= N) | (c1_ >= N)
        outer0 = c0_ >= N
        l0 = np.where(c0_ < N, local(t, np.minimum(c0_, N - 1)), LOCAL_MAX)
        l1 = np.where(c1_ < N, local(t, np.minimum(c1_, N - 1)), LOCAL_MAX)
        if np.any(l0 < 0) or np.any(l1 < 0):
            raise AssertionError("facet side not staged")
        k0 = np.where(c0_ < N, side_slot[0, f], 0)
        k1 = np.where(c1_ < N, side_slot[1, f], 0)
        lf = (c0_ < c1_).astype(np.int64)
        d = (l0 | (k0 << 12) | (l1 << 14) | (k1 << 26) | (lf << 28) | (bnd.astype(np.int64) << 29)
             | (outer0.astype(np.int64) << 30))
        return d.astype(np.uint32)

    fdesc = descriptors(own_t, np.arange(F))
    fdesc_x = descriptors(ft, fx)
    # compact facet geometry: n [D], 1/(n_hat.d), X_orthog's tangent coordinates (2 D
    # floats; the kernel derives |n| and n_hat from n as the reference's FacetCalc does,
    # facet_process.py:84-86, and X_orthog from n_hat)
    geo_c = compact_geometry(geo, D)
    geo_x = geo_c[fx]

    # ---- tile-local facet index of each cell slot ----
    cf_t = np.repeat(tile_of, NL)
    cf_f = cell_facet.reshape(-1)
    tl = np.where(own_t[cf_f] == cf_t, cf_f - fo0[cf_t], -1)
    xk = ft * F + fx                                   # sorted (ft, fx)
    need = tl < 0
    px = _lookup(xk, cf_t[need] * F + cf_f[need])
    if np.any(px < 0):
        raise AssertionError("cell facet neither owned nor foreign")
    tl[need] = nfo[cf_t[need]] + px - foff[cf_t[need]]
    cfac = np.full((N, 4), 0, np.uint16)
    cfac[:, :NL] = tl.reshape(N, NL).astype(np.uint16)

    # ---- H1 facet table: for each slot of each H1 cell, the tile-local position of the
    # facet if it is in the tile's list (| SIDE_BIT when the H1 cell is its slot-1 side),
    # else NONE16: the H1 cell's gradient thread writes its face value there ----
    h1f = np.full((len(K1), 4), NONE16, np.uint16)
    if len(K1):
        hf_t = np.repeat(t1, NL)
        hf_f = cell_facet[h1c].reshape(-1)
        pos = np.where(own_t[hf_f] == hf_t, hf_f - fo0[hf_t], -1)
        rest = pos < 0
        px = _lookup(xk, hf_t[rest] * F + hf_f[rest])
        pos[rest] = np.where(px >= 0, nfo[hf_t[rest]] + px - foff[hf_t[rest]], -1)
        side1 = fac_cell[1, hf_f] == np.repeat(h1c, NL)
        val = np.where(pos >= 0, pos | np.where(side1, SIDE_BIT, 0), NONE16)
        h1f[:, :NL] = val.reshape(-1, NL).astype(np.uint16)

    # ---- characteristic-group means a tile needs (groups of the meshes it touches) ----
    gcell = np.asarray(bcell, np.int64)[np.asarray(group) >= 0]
    needs = np.zeros(T, bool)
    if len(gcell):
        has = np.zeros(N, bool)
        has[gcell] = True
        needs[tile_of[gcell]] = True
        needs[t1[has[h1c]]] = True
    cm = np.asarray(cell_mesh, np.int64)
    m_lo = cm[np.arange(T) * TILE]
    m_hi = cm[np.minimum(np.arange(T) * TILE + TILE - 1, N - 1)]
    g0 = np.asarray(group_off, np.int64)[m_lo]
    ng = np.where(needs, np.asarray(group_off, np.int64)[m_hi + 1] - g0, 0)
    if np.any(ng > 4):
        raise TilesUnavailable("a tile touches more than 4 boundary groups")

    tab = np.zeros((T, TAB), np.int32)
    halo_off = np.concatenate([[0], np.cumsum(n1 + n2)])
    tab[:, 0] = halo_off[:-1]
    tab[:, 1] = n1
    tab[:, 2] = n2
    tab[:, 3] = nbr_off[:-1]
    tab[:, 4] = fo0
    tab[:, 5] = nfo
    tab[:, 6] = foff[:-1]
    tab[:, 7] = nfor
    tab[:, 8] = g0
    tab[:, 9] = ng
    tab[:, 10] = s1[:-1]
    halo = np.zeros(int(halo_off[-1]), np.int32)
    # H1 then H2 of each tile
    halo[halo_off[t1] + (np.arange(len(K1)) - s1[t1])] = h1c
    halo[halo_off[t2] + n1[t2] + (np.arange(len(K2)) - s2[t2])] = K2 % N

    # ---- packed slot words of the kernel: tile cells cw [NL][N] (global cell index), H1
    # entries hw [NL][sum n1] (H1 entry index): see pack_slot_words ----
    if gref is None:
        gref = np.broadcast_to(np.arange(NL, dtype=np.int64), (N, NL))
    gref = np.asarray(gref, np.int64)
    own_side = (fac_cell[1, cell_facet] == np.arange(N)[:, None]).astype(np.int64)
    own_nl = rows[own_row, :NL].astype(np.int64)
    cw = pack_slot_words(np.where(own_nl == NONE16, NONE12, own_nl), cfac[:, :NL].astype(np.int64), own_side,
                         gref)
    hv = h1f[:, :NL].astype(np.int64)
    h_nl = rows[h1_row, :NL].astype(np.int64)
    hw = pack_slot_words(np.where(h_nl == NONE16, NONE12, h_nl), np.where(hv == NONE16, NONE10, hv & ~SIDE_BIT),
                         np.where(hv == NONE16, 0, (hv & SIDE_BIT) != 0).astype(np.int64), gref[h1c])

    fpt = facets_per_thread(D)
    if int(n1.max(initial=0)) > TILE or int((nfo + nfor).max(initial=0)) > fpt * TILE:
        raise TilesUnavailable(f"a tile has more than {TILE} halo cells or {fpt * TILE} facets")
    ncl = TILE + int((n1 + n2).max(initial=0))
    ngl = TILE + int(n1.max(initial=0))
    nfl = int((nfo + nfor).max(initial=1))
    batch = int(np.max(cm, initial=0)) > 0
    lds = fused_lds_bytes(D, ncl, ngl, nfl, batch)
    if lds > lds_limit:
        raise TilesUnavailable(f"tile working set {lds} B exceeds {lds_limit} B of LDS")
    if nfl > NONE10:
        raise TilesUnavailable(f"a tile lists more than {NONE10} facets")
    return dict(tab=tab, halo=halo, hrec=hrec, nbr=rows, h1f=h1f, cw=np.ascontiguousarray(cw.T),
                hw=np.ascontiguousarray(hw.T), fdesc=fdesc, ffac=fx.astype(np.int32),
                fdesc_x=fdesc_x, geo=geo_c, geo_x=geo_x, cfac=cfac, ncl=ncl, ngl=ngl, nfl=nfl, nh1=len(K1),
                lds_bytes=lds,
                lds_fits_target=None if lds_target is None else lds <= lds_target, n_tiles=T,
                halo_cells=(int(n1.sum()), int(n2.sum())), foreign=int(nfor.sum()))


def check_tiles(t: dict, N: int, D: int, nbr, cell_facet, fac_cell, gref=None) -> None:
    """Rebuild, from the tile tables alone, every global relation the fused kernel relies
    on and compare it with the mesh (raises AssertionError on the first mismatch).  The
    packed slot words (what the kernel reads) must decode to the same neighbour rows,
    list positions and sides as the unpacked tables, and to the slots' reference order."""
    NL = D + 1
    nbr = np.asarray(nbr, np.int64)
    cell_facet = np.asarray(cell_facet, np.int64)
    fac_cell = np.asarray(fac_cell, np.int64)
    tab, halo = t["tab"], t["halo"]
    # packed words of the tile cells against the unpacked tables
    nl, q, side, ref = unpack_slot_words(t["cw"].T)
    orow = t["nbr"][(np.repeat(tab[:, 3], TILE) + np.tile(np.arange(TILE), t["n_tiles"]))[:N], :NL].astype(np.int64)
    assert np.array_equal(np.where(orow == NONE16, NONE12, orow), nl)
    assert np.array_equal(t["cfac"][:, :NL].astype(np.int64), q)
    assert np.array_equal(side, (fac_cell[1, cell_facet] == np.arange(N)[:, None]).astype(np.int64))
    if gref is not None:
        assert np.array_equal(np.asarray(gref, np.int64), ref)
    hn, hq, hs, _ = unpack_slot_words(t["hw"].T)
    hv = t["h1f"][:, :NL].astype(np.int64)
    assert np.array_equal(np.where(hv == NONE16, NONE10, hv & ~SIDE_BIT), hq)
    assert np.array_equal(np.where(hv == NONE16, 0, (hv & SIDE_BIT) != 0), hs)
    for ti in range(t["n_tiles"]):
        ho, n1, n2, no, fo0, nfo, foff, nfor = (int(v) for v in tab[ti, :8])
        c0 = ti * TILE
        nown = min(TILE, N - c0)
        glob = np.full(TILE + n1 + n2, -1, np.int64)
        glob[:nown] = c0 + np.arange(nown)
        glob[TILE:] = halo[ho:ho + n1 + n2]
        rows = t["nbr"][no:no + TILE + n1].astype(np.int64)
        for l in list(range(nown)) + list(range(TILE, TILE + n1)):
            g = glob[l]
            for k in range(NL):
                r = rows[l, k]
                want = nbr[g, k]
                assert (r == NONE16) == (want >= N), (ti, l, k)
                if r != NONE16:
                    assert glob[r] == want, (ti, l, k, glob[r], want)
        fl = np.concatenate([np.arange(fo0, fo0 + nfo), t["ffac"][foff:foff + nfor]])
        dl = np.concatenate([t["fdesc"][fo0:fo0 + nfo], t["fdesc_x"][foff:foff + nfor]]).astype(np.int64)
        for f, d in zip(fl, dl):
            for s, (sh_l, sh_k) in enumerate(((0, 12), (14, 26))):
                l = (d >> sh_l) & LOCAL_MAX
                k = (d >> sh_k) & 3
                c = fac_cell[s, f]
                if c >= N:
                    assert l == LOCAL_MAX and ((d >> 29) & 1) and ((d >> 30) & 1) == (s == 0)
                else:
                    assert glob[l] == c and cell_facet[c, k] == f, (ti, f, s)
            assert ((d >> 28) & 1) == (fac_cell[0, f] < fac_cell[1, f])
        for i in range(nown):
            c = c0 + i
            for k in range(NL):
                assert fl[t["cfac"][c, k]] == cell_facet[c, k], (ti, c, k)
        # H1 facet table: every (facet of the list, side) is written by exactly one cell
        seen = set()
        for i in range(nown):
            c = c0 + i
            for k in range(NL):
                q = int(t["cfac"][c, k])
                seen.add((q, int(fac_cell[1, fl[q]] == c)))
        for j in range(n1):
            h = glob[TILE + j]
            for k in range(NL):
                v = int(t["h1f"][int(tab[ti, 10]) + j, k])
                f = cell_facet[h, k]
                if v == NONE16:
                    assert f not in set(fl.tolist()), (ti, h, k)
                    continue
                q, side = v & ~SIDE_BIT, int(bool(v & SIDE_BIT))
                assert fl[q] == f and fac_cell[side, f] == h, (ti, h, k)
                assert (q, side) not in seen
                seen.add((q, side))
        for q, f in enumerate(fl):
            for side in (0, 1):
                if fac_cell[side, f] < N:
                    assert (q, side) in seen, (ti, q, side)
