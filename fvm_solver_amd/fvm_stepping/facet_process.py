"""Device mesh + fused RHS driver -- replaces ``time_fvm/fvm_stepping/facet_process.py``.

``FacetCalc`` keeps the reference's constructor (facet_process.py:261) and role: it owns
the device-resident mesh operators and evaluates the face/flux pipeline.  Instead of the
reference's SPM objects and ~150 ATen launches per RHS, the operators are packed into the
per-cell / per-facet arrays of ``FvmMeshDev`` (include/fvm_hip.h) and one RHS is four
HIP launches (see csrc/rhs.hip).  What the reference assembles as

    flux_mat    (facet_process.py:200-226)   -> cell_inv_vol + slot bits, facets ascending
    G_mats      (fvm_mesh.py:5-42, 263-344)  -> cell_gval in slot order (neighbours, then the cell)
    A_face_grad (facet_process.py:114-198)   -> fac_geo[1/(n.d)] + bc_ga/bc_gb
    A_bc, b_bc  (facet_boundary.py:238-280)  -> bc_flags + bc_val
    A_corr      (facet_boundary.py:170-236)  -> bc_l
    BC groups   (facet_boundary.py:32-80, 289-385) -> bc_group + bc_char

holds the same numbers (float32 values rounded from the float64 geometry).  State kept
between calls: the unlimited gradient and limiter factor of the last call (``work.gp``),
from which ``cell_grads`` (= the reference's limited gradients) is derived on request.
"""
from __future__ import annotations

import numpy as np
import torch

from fvm_solver_amd import _native as nat
from fvm_solver_amd.config_fvm import BCMode
from fvm_solver_amd.fvm_stepping.fused_tiles import TilesUnavailable, build_tiles, device_geometry

_K_BC_PART = 64


def _f32(x):
    return np.asarray(x, dtype=np.float32)


def _bc_tables(bc_tags, C):
    """Flags and values per boundary facet and component from the Facet descriptors
    (FacetCalc._init_bc, facet_process.py:297-330)."""
    nb = len(bc_tags)
    dm = np.zeros((nb, C), bool)
    nm = np.zeros((nb, C), bool)
    dv = np.full((nb, C), np.nan, np.float32)
    nv = np.full((nb, C), np.nan, np.float32)
    ff = np.zeros(nb, bool)
    il = np.zeros(nb, bool)
    for i, (k, e) in enumerate(bc_tags.items()):
        if k != i:
            raise ValueError("bc_tags must be keyed 0..F_b-1 in boundary-facet order")
        dm[i] = e.dirichlet()
        nm[i] = e.neumann()
        dv[i] = np.asarray(e.U, np.float32)
        nv[i] = np.asarray(e.dUdn, np.float32)
        ff[i] = e.farfield()
        il[i] = e.inlet()
    return dm, nm, dv, nv, ff, il


def host_layout(mesh, n_comp: int, bc_tags, cfg) -> dict:
    """Per-mesh FvmMeshDev arrays (numpy, mesh-local indices) from an FVMMesh and its BC
    tags; ``DeviceMesh`` concatenates one or more of these and uploads them."""
    g = mesh.geo64
    D = mesh.dim
    C = n_comp
    if C != D + 2:
        raise ValueError(f"n_comp must be dim + 2 (got {C} for dim {D})")
    NL = D + 1
    N, F, Fb = mesh.n_cells, mesh.n_facets, mesh.n_bc_facet
    if N >= 2 ** 31 - 1 or F >= 2 ** 31 - 1:
        raise ValueError("mesh too large for int32 indices")
    bcm = mesh.bc_facet_mask.numpy()
    c2f, signs, comb = g["c2f"], g["signs"], g["combined"]
    slots = (signs == -1).astype(np.int64)           # int(-s + 1/2), facet_process.py:71

    # ---- per-cell lists in ascending facet order (flux_mat ELL order) ----
    perm = np.argsort(c2f, axis=1, kind="stable")
    c2f_s = np.take_along_axis(c2f, perm, 1)
    nbr_s = np.take_along_axis(comb, perm, 1)
    slot_s = np.take_along_axis(slots, perm, 1)
    # centroid -> facet-midpoint displacements in slot order, fp32; a cell's displacements
    # sum to zero (the facet midpoints average to the centroid), so the last slot's is
    # stored as minus the fp32 sum of the others: the tile-fused stage derives it instead
    # of loading it, and every kernel then sees the same floats
    disp_s = np.take_along_axis(g["cent_to_facet_disp"], perm[..., None], 1).astype(np.float32)
    dsum = disp_s[:, 0, :].copy()
    for k in range(1, NL - 1):
        dsum = (dsum + disp_s[:, k, :]).astype(np.float32)
    disp_s[:, NL - 1, :] = -dsum
    flat = c2f_s.reshape(-1)
    o = np.argsort(flat, kind="stable")
    counts = np.bincount(flat, minlength=F)
    start = np.concatenate([[0], np.cumsum(counts)[:-1]])
    two = counts == 2
    pa = o[start[two]]
    pb = o[start[two] + 1]
    nbk = np.zeros(N * NL, np.int64)
    nbk[pa] = pb % NL
    nbk[pb] = pa % NL
    nbk = nbk.reshape(N, NL)

    # ---- LSQ rows (G_mats' values) in slot order: the neighbour in each slot, then the
    # cell itself (the kernels sum the row in this order) ----
    G32 = g["G"].astype(np.float32)                   # [N, D, NL], reference local order
    diag = G32[:, :, 0].copy()
    for k in range(1, NL):
        diag = (diag + G32[:, :, k]).astype(np.float32)
    gval = np.concatenate([np.take_along_axis(G32, perm[:, None, :], 2), -diag[:, :, None]], 2)  # [N, D, NL+1]

    info = np.zeros(N, np.int64)
    for k in range(NL):
        info |= slot_s[:, k] << k
        info |= nbk[:, k] << (8 + 2 * k)

    vol32 = g["volumes"].astype(np.float32)
    inv_vol = (np.float32(1.0) / vol32).astype(np.float32)
    eps = np.power(np.float32(cfg.lim_K) * np.sqrt(vol32), np.float32(cfg.lim_p + 1)).astype(np.float32)

    # ---- per-facet geometry ----
    n64, nh64 = g["normals"], g["normals_hat"]
    fs64 = np.linalg.norm(n64, axis=1)
    cen = g["centroids"]
    GEO = 12 if D == 2 else 20
    geo = np.zeros((F, GEO), np.float32)
    geo[:, :D] = n64
    geo[:, D:2 * D] = nh64
    geo[:, 2 * D] = fs64
    inter = ~bcm
    cd = cen[g["right"][inter]] - cen[g["left"][inter]]
    dproj = np.einsum("nd,nd->n", nh64[inter], cd)
    geo[inter, 2 * D + 1] = 1.0 / dproj
    geo[inter, 2 * D + 2:3 * D + 2] = cd / dproj[:, None] - nh64[inter]
    # centroid -> facet-midpoint displacement of the slot-0 and slot-1 cells (the cells'
    # own fp32 displacement lists, above)
    for k in range(NL):
        f_k = c2f_s[:, k]
        sl = slot_s[:, k]
        for sv in (0, 1):
            sel = sl == sv
            geo[f_k[sel], 3 * D + 2 + sv * D:3 * D + 2 + (sv + 1) * D] = disp_s[sel, k]

    # ---- boundary facets ----
    dm, nm, dv, nv, ffm, ilm = _bc_tables(bc_tags, C)
    if len(dm) != Fb:
        raise ValueError(f"{len(dm)} BC tags for {Fb} boundary facets")
    fidx = np.flatnonzero(bcm)
    bcell = g["left"][bcm]
    kloc = np.argmax(c2f[bcell] == fidx[:, None], axis=1)
    cell_slot = slots[bcell, kloc]
    side = 1 - cell_slot                                   # the free slot (facet_process.py:98-101)
    mid = g["midpoints"][bcm]
    dist64 = np.einsum("nd,nd->n", mid - cen[bcell], nh64[bcm])
    dist32 = dist64.astype(np.float32)
    nv_masked = nv[nm]                                      # Neumann values, row-major entry order
    comp = np.broadcast_to(np.arange(C), (Fb, C))
    bc_val = np.zeros((Fb, C), np.float32)
    bc_val[dm] = dv[dm]
    bc_ga = np.zeros((Fb, C), np.float32)
    bc_gb = np.zeros((Fb, C), np.float32)
    if dm.any():
        dd = np.broadcast_to(dist32[:, None], (Fb, C))[dm]
        bc_ga[dm] = np.float32(-1.0) / dd
        bc_gb[dm] = dv[dm] / dd
    if nm.any():
        # b_bc and b_face_grad index the masked Neumann values by component
        # (facet_boundary.py:278, facet_process.py:192); kept as the reference does it.
        if comp[nm].max() >= len(nv_masked):
            raise IndexError("Neumann values indexed by component out of range")
        bc_val[nm] = nv_masked[comp[nm]] * np.broadcast_to(dist32[:, None], (Fb, C))[nm]
        bc_gb[nm] = nv_masked[comp[nm]]
    flags = (nm * (1 << np.arange(C))).sum(1).astype(np.int32)
    r = g["cent_to_facet_disp"][bcell, kloc]
    nhb = nh64[bcm]
    l_vec = r - nhb * np.einsum("nd,nd->n", r, nhb)[:, None]
    l_vec[~nm.any(1)] = 0.0

    # ---- characteristic BC groups (facet_boundary.py:32-80, 289-385) ----
    nh32 = nhb.astype(np.float32)
    sgn = (np.float32(2.0) * (side.astype(np.float32) - np.float32(0.5))).astype(np.float32)
    n_out = (nh32 * sgn[:, None]).astype(np.float32)
    group = np.full(Fb, -1, np.int32)
    CH = 2 * D + 3
    bc_char = np.zeros((Fb, CH), np.float32)
    modes = []
    R32 = np.float32(cfg.C_v * (cfg.gamma - 1))
    for mask, bcfg in ((ffm, cfg.exit_cfg), (ilm, cfg.inlet_cfg)):
        if not mask.any():
            continue
        if bcfg.mode != BCMode.Characteristic:
            raise NotImplementedError(f"Unknown BCMode {bcfg.mode}: only the characteristic BC is supported "
                                      "(the reference's Farfield mode is broken and Isentropic is unimplemented)")
        if bcfg.mode not in modes:
            modes.append(bcfg.mode)
        group[mask] = modes.index(bcfg.mode)
        nn = n_out[mask]
        rho = np.float32(bcfg.rho_inf)
        T = np.float32(bcfg.T_inf)
        bc_char[mask, :D] = nn
        bc_char[mask, D] = rho
        bc_char[mask, D + 2] = (R32 * rho) * T
        if D == 2:
            bc_char[mask, D + 1] = np.float32(-bcfg.v_n_inf)
            tang = np.stack([-nn[:, 1], nn[:, 0]], 1)
            bc_char[mask, D + 3:] = tang * np.float32(bcfg.v_t_inf)
        else:
            vinf = _f32(bcfg.v_inf)
            vn = (nn[:, 0] * vinf[0] + nn[:, 1] * vinf[1]) + nn[:, 2] * vinf[2]
            bc_char[mask, D + 1] = vn
            bc_char[mask, D + 3:] = vinf[None, :] - nn * vn[:, None]
    per = fs64[c2f].sum(1)
    return dict(D=D, C=C, N=N, F=F, Fb=Fb, n_groups=len(modes), c2f_s=c2f_s, nbr_s=nbr_s, slot_s=slot_s, nbk=nbk,
                gval=gval, gref=perm.astype(np.int8), info=info, disp_s=disp_s, inv_vol=inv_vol, eps=eps, geo=geo,
                bcell=bcell,
                group=group, flags=flags, bc_val=bc_val, bc_l=l_vec.astype(np.float32), bc_ga=bc_ga, bc_gb=bc_gb,
                bc_char=bc_char, side=side, volumes64=g["volumes"], h64=D * g["volumes"] / per,
                phys=(float(cfg.viscosity), float(cfg.visc_bulk), float(cfg.thermal_cond)))


_TILE = 256


class DeviceMesh:
    """The FvmMeshDev arrays of one mesh, or of several meshes concatenated
    block-diagonally (``DeviceMesh.batch``), uploaded to ``device``.

    Meshes are not padded to 256-cell tiles: a tile may hold cells of two meshes, and
    the boundary-mean slot tables list, per tile, the groups of every mesh the tile
    touches (at most 4)."""

    def __init__(self, mesh, n_comp: int, bc_tags, cfg, device, persist_blocks: int = 0, tiles: bool = True,
                 lds_limit: int | None = None):
        self._build([host_layout(mesh, n_comp, bc_tags, cfg)], device, persist_blocks, tiles, lds_limit)

    @classmethod
    def batch(cls, items, device, persist_blocks: int = 0, tiles: bool = True, lds_limit: int | None = None):
        """items: iterable of (mesh, n_comp, bc_tags, cfg)."""
        obj = cls.__new__(cls)
        obj._build([host_layout(*it) for it in items], device, persist_blocks, tiles, lds_limit)
        return obj

    @staticmethod
    def lds_limit_for(device, D: int) -> int:
        """LDS bytes one workgroup of the fused kernel may use: the device's limit per
        workgroup (160 KiB on gfx950), or, for a host-only layout, 64 KiB in 2-D (what
        lets several tiles share a CU) and 160 KiB in 3-D."""
        dev = torch.device(device)
        if dev.type == "cuda":
            return int(nat.max_lds_per_workgroup(dev))
        return 64 * 1024 if D == 2 else 160 * 1024

    def _build(self, L, device, persist_blocks, want_tiles=True, lds_limit=None):
        D, C = L[0]["D"], L[0]["C"]
        if any(l["D"] != D for l in L):
            raise ValueError("all meshes of a batch must have the same dimension")
        NL = D + 1
        nm = len(L)
        N = sum(l["N"] for l in L)
        F = sum(l["F"] for l in L)
        Fb = sum(l["Fb"] for l in L)
        if N >= 2 ** 30 or F >= 2 ** 31 - 1:
            raise ValueError("mesh too large for int32 indices")
        self.dim, self.C, self.N, self.F, self.Fb, self.n_meshes = D, C, N, F, Fb, nm
        co = np.cumsum([0] + [l["N"] for l in L])
        fo = np.cumsum([0] + [l["F"] for l in L])
        bo = np.cumsum([0] + [l["Fb"] for l in L])
        go = np.cumsum([0] + [l["n_groups"] for l in L])
        self.cell_offsets, self.facet_offsets, self.bc_offsets, self.group_offsets = co, fo, bo, go
        self.n_groups = int(go[-1])

        def cat(key, fn=lambda a, i: a):
            return np.concatenate([fn(l[key], i) for i, l in enumerate(L)])

        c2f_s = cat("c2f_s", lambda a, i: a + fo[i])
        nbr_s = cat("nbr_s", lambda a, i: np.where(a < L[i]["N"], a + co[i], a - L[i]["N"] + bo[i] + N))
        bcell = cat("bcell", lambda a, i: a + co[i])
        group = cat("group", lambda a, i: np.where(a >= 0, a + go[i], -1)).astype(np.int32)
        cell_mesh = np.repeat(np.arange(nm, dtype=np.int32), [l["N"] for l in L])

        # boundary-mean slots: tiles holding the interior cells of group facets
        ntiles = (N + _TILE - 1) // _TILE
        tiles = np.unique(bcell[group >= 0] // _TILE) if self.n_groups else np.zeros(0, np.int64)
        blk_slot = np.full(ntiles, -1, np.int32)
        blk_slot[tiles] = np.arange(len(tiles), dtype=np.int32)
        m_lo = cell_mesh[tiles * _TILE] if len(tiles) else np.zeros(0, np.int32)
        m_hi = cell_mesh[np.minimum(tiles * _TILE + _TILE - 1, N - 1)] if len(tiles) else np.zeros(0, np.int32)
        slot_g0 = go[m_lo].astype(np.int32)
        slot_ng = (go[m_hi + 1] - go[m_lo]).astype(np.int32)
        if len(tiles) and slot_ng.max() > 4:
            raise ValueError("a 256-cell tile touches more than 4 boundary groups (meshes too small to batch)")
        gmesh = np.repeat(np.arange(nm), [l["n_groups"] for l in L])
        group_s0 = np.searchsorted(m_hi, gmesh, side="left").astype(np.int32)
        group_s1 = np.searchsorted(m_lo, gmesh, side="right").astype(np.int32)
        group_count = np.array([np.sum(group == q) for q in range(self.n_groups)], np.float32)
        self.n_slots = len(tiles)
        self.host = dict(c2f_s=c2f_s, nbr_s=nbr_s, bcell=bcell, group=group, cell_mesh=cell_mesh,
                         blk_slot=blk_slot, slot_g0=slot_g0, slot_ng=slot_ng, group_s0=group_s0,
                         group_s1=group_s1, group_count=group_count)
        for k in ("slot_s", "nbk", "gval", "gref", "info", "disp_s", "inv_vol", "eps", "geo", "flags", "bc_val",
                  "bc_l", "bc_ga", "bc_gb", "bc_char", "side"):
            self.host[k] = cat(k)
        h = self.host
        self.mesh_phys = np.array([[*l["phys"], 0.0] for l in L], np.float32)

        # the two sides of every facet, and the device facet order: by lowest adjacent
        # cell, so a run of facets touches a run of cells (cells are space-filling-curve
        # ordered) and each cell finds its facets' fluxes close together; within the run
        # of a 256-cell tile the boundary facets go last, so the waves that evaluate a
        # tile's facets meet the boundary-condition branch in one place (every facet's
        # flux and every cell's sum over its slots are the same in any facet order)
        slot_s = h["slot_s"]
        fac_cell = np.full((2, F), -1, np.int64)
        fac_cell[slot_s, c2f_s] = np.broadcast_to(np.arange(N)[:, None], c2f_s.shape)
        outer = nbr_s >= N
        fac_cell[1 - slot_s[outer], c2f_s[outer]] = nbr_s[outer]
        if (fac_cell < 0).any():
            raise ValueError("a facet is missing one of its two sides")
        lo = fac_cell.min(0)
        fperm = np.lexsort((lo, fac_cell.max(0) >= N, lo // 256))
        fpos = np.empty(F, np.int64)
        fpos[fperm] = np.arange(F)
        h["fac_cell"], h["fperm"], h["fpos"] = fac_cell, fperm, fpos

        # tables of the tile-fused stage kernel
        self.tiles, self.tiles_reason = None, "not built" if want_tiles else "disabled"
        if want_tiles:
            lim = lds_limit if lds_limit is not None else self.lds_limit_for(device, D)
            try:
                self.tiles = build_tiles(N, D, nbr_s, fpos[c2f_s], fac_cell[:, fperm], h["gval"], h["disp_s"],
                                         h["eps"], h["geo"][fperm], bcell, group, cell_mesh, go, lds_limit=lim,
                                         gref=h["gref"])
                self.tiles_reason = None
            except TilesUnavailable as e:
                self.tiles_reason = str(e)

        dev = torch.device(device)

        def up(a, dt):
            return torch.from_numpy(np.ascontiguousarray(a)).to(dtype=dt, device=dev)

        def up_nz(a, dt):
            return up(a if len(a) else np.zeros(1), dt)

        i32, f32 = torch.int32, torch.float32
        self.t = dict(
            cell_facet=up(fpos[c2f_s].T, i32), cell_nbr=up(nbr_s.T, i32), cell_info=up(h["info"], i32),
            cell_disp=up(h["disp_s"].transpose(1, 2, 0), f32), cell_gval=up(h["gval"].transpose(2, 1, 0), f32),
            cell_inv_vol=up(h["inv_vol"], f32), cell_eps=up(h["eps"], f32), fac_geo=up(device_geometry(h["geo"], D)[fperm], f32),
            fac_cell=up(fac_cell[:, fperm], i32),
            bc_cell=up_nz(bcell, i32), bc_group=up_nz(group, i32), bc_flags=up_nz(h["flags"], i32),
            bc_val=up_nz(h["bc_val"], f32), bc_l=up_nz(h["bc_l"], f32), bc_ga=up_nz(h["bc_ga"], f32),
            bc_gb=up_nz(h["bc_gb"], f32), bc_char=up_nz(h["bc_char"], f32),
            blk_slot=up(blk_slot, i32), slot_blk=up_nz(tiles, i32), slot_g0=up_nz(slot_g0, i32),
            slot_ng=up_nz(slot_ng, i32), group_s0=up_nz(group_s0, i32), group_s1=up_nz(group_s1, i32),
            group_count=up_nz(group_count, f32))
        if nm > 1:
            self.t["cell_mesh"] = up(cell_mesh, i32)
            self.t["mesh_phys"] = up(self.mesh_phys, f32)
        self.bc_facet_side = h["side"]
        self.bc_cell_np = bcell
        if dev.type != "cuda":      # host-only layout (inspection and tests)
            self.t, self.struct, self.struct3 = None, None, None
            return
        self.volumes64 = up(cat("volumes64"), torch.float64)
        self.h64 = up(cat("h64"), torch.float64)
        base = dict(dim=D, n_groups=self.n_groups, n_cells=N, n_facets=F, n_bc=Fb, n_slots=self.n_slots, n_meshes=nm,
                    persist_blocks=int(persist_blocks), **{k: v.data_ptr() for k, v in self.t.items()})
        # struct3: the three-kernel pipeline; struct: the tile-fused stage when tiles exist
        self.struct3 = nat.FvmMeshDev(**base)
        if self.tiles is None:
            self.struct = self.struct3
            return
        tl = self.tiles
        self.tz = dict(
            fz_tile=up(tl["tab"], i32), fz_halo=up_nz(tl["halo"], i32),
            fz_hrec=up(tl["hrec"] if tl["nh1"] else np.zeros(1), f32),
            fz_cw=up(tl["cw"].view(np.int32), i32), fz_hw=up_nz(tl["hw"].view(np.int32).reshape(-1), i32),
            fz_fdesc=up(tl["fdesc"].view(np.int32), i32),
            fz_ffac=up_nz(tl["ffac"], i32), fz_fdesc_x=up_nz(tl["fdesc_x"].view(np.int32), i32),
            fz_geo=up(tl["geo"], f32), fz_geo_x=up_nz(tl["geo_x"], f32))
        self.struct = nat.FvmMeshDev(**base, fz_ncl=tl["ncl"], fz_ngl=tl["ngl"], fz_nfl=tl["nfl"], fz_nh1=tl["nh1"],
                                     **{k: v.data_ptr() for k, v in self.tz.items()})

    def set_persist_blocks(self, n: int):
        """Workgroups of the per-cell kernels of the three-kernel pipeline (0 = one per
        256-cell tile); must be a multiple of 8 so the XCD-aware tile mapping stays a
        bijection.  The tile-fused stage always runs one workgroup per tile."""
        if n < 0 or n % 8:
            raise ValueError("persist_blocks must be a non-negative multiple of 8")
        self.struct.persist_blocks = int(n)
        self.struct3.persist_blocks = int(n)

    def fused_thread_options(self) -> list:
        """Workgroup sizes the tile-fused stage can run this mesh with: 256 (a thread per
        tile cell, which also takes an H1 cell and up to 2 / 3 facets), and 320, 384, 512
        where the largest tile's H1 fits the threads beyond the first 256 (tile and H1
        cells on separate threads, fewer facets each).  All compute bitwise the same."""
        if self.tiles is None:
            return []
        return [256] + [n for n in (320, 384, 512) if self.tiles["ngl"] <= n]

    def set_fused_threads(self, n: int):
        if n not in self.fused_thread_options():
            raise ValueError(f"fused-stage threads must be one of {self.fused_thread_options()}")
        self.struct.fz_threads = int(n)

    @property
    def fused_threads(self) -> int:
        return int(self.struct.fz_threads) or 256 if self.tiles is not None else 0


def phys_struct(cfg) -> "nat.FvmPhys":
    """Scalars of FluidConstitution / SlopeLimiter / KTDiffusion, rounded to float32 from
    the double-precision Python expressions the reference evaluates them with."""
    R = cfg.C_v * (cfg.gamma - 1)
    return nat.FvmPhys(inv_cv=1 / cfg.C_v, cv=cfg.C_v, R=R, gR=cfg.gamma * R, mu0=cfg.viscosity, T0=cfg.T_0,
                       T0pS=cfg.T_0 + cfg.S_const, S=cfg.S_const, mub0=cfg.visc_bulk, T0sq=cfg.T_0 ** 2,
                       kappa=cfg.thermal_cond, v_factor=cfg.v_factor, lim_p=int(cfg.lim_p))


class FacetCalc:
    """Face/flux pipeline owner (facet_process.py:229-450) on the GPU."""

    def __init__(self, phy_setup, cfg, mesh_setup, n_comp: int, bc_tags, device="cuda"):
        device = torch.device(device)
        if device.type != "cuda":
            raise RuntimeError("FacetCalc runs on a GPU device only")
        nat.load_library()
        self.phy_setup = phy_setup
        self.fvm_mesh = mesh_setup
        self.bc_type_str = [e.tag for e in bc_tags.values()]
        self._init(DeviceMesh(mesh_setup, n_comp, bc_tags, cfg, device), cfg, device)
        self.fvm_mesh = mesh_setup            # the reference keeps it as E_props.mesh.fvm_mesh
        self.mesh.fvm_mesh = mesh_setup
        self.mesh.bc_facet_mask = mesh_setup.bc_facet_mask

    @classmethod
    def from_device_mesh(cls, dm: "DeviceMesh", cfg, device="cuda", fused: bool = True):
        """A FacetCalc over an already built (e.g. batched) device layout.  fused=False
        keeps the three-kernel pipeline even where tile tables exist (callers that read
        every cell's gradient record after each call)."""
        obj = cls.__new__(cls)
        nat.load_library()
        obj.phy_setup = None
        obj.fvm_mesh = None
        obj.bc_type_str = []
        obj._init(dm, cfg, torch.device(device), fused)
        return obj

    def _init(self, dm, cfg, device, fused=True):
        self.device = device
        self.cfg = cfg
        self.mesh = dm
        self.n_comp = dm.C
        self.dim = dm.dim
        self.n_local = self.dim + 1
        self.n_cells = dm.N
        self.n_facets = dm.F
        self.n_facets_bc = dm.Fb
        self.phys = phys_struct(cfg)
        D, C, N, Fb = self.dim, self.n_comp, self.n_cells, self.n_facets_bc
        dev = self.device
        self.nblk = (N + 255) // 256
        self.ubc = torch.zeros((max(Fb, 1), C), dtype=torch.float32, device=dev)
        # gradient records: updated in place by the three-kernel pipeline; the fused
        # stage reads the previous call's records of boundary cells while writing the
        # new ones, so with tile tables there is a second buffer and the two swap
        self._gpbuf = [torch.zeros((N, (D + 1) * C), dtype=torch.float32, device=dev)
                       for _ in range(2 if dm.tiles is not None else 1)]
        self._latest = 0             # buffer holding the last call's records
        self._gfull = None
        # boundary-mean partials: a stage reads one buffer (the sums of its input) and its
        # epilogue writes the other (the sums of the next stage's input); they swap after
        # every stage
        self._bcp = [torch.zeros(4 * max(self.mesh.n_slots, 1), dtype=torch.float32, device=dev) for _ in range(2)]
        self._bcp_cur = 0
        self.red = torch.zeros(self.nblk * 3 * C, dtype=torch.float64, device=dev)
        self.fflux = torch.zeros((self.n_facets, C), dtype=torch.float32, device=dev)
        self.set_fused(fused)
        self.work = nat.FvmWork(ubc=self.ubc.data_ptr(), gp=self._gpbuf[0].data_ptr(), bc_part=self._bcp[0].data_ptr(),
                                red=self.red.data_ptr(), fflux=self.fflux.data_ptr(), gp_prev=None,
                                bc_part_next=self._bcp[1].data_ptr())
        self.has_prev = False
        self._last_has_prev = False
        self._last_fused = False
        self.bc_ready_ptr = None     # prim buffer whose boundary sums bc_part holds
        self.last_prim = None        # primitives of the last evaluation, if known (saving)

    @property
    def bc_part(self):
        """Boundary-mean partials of the next stage's input (read by its gradient phase)."""
        return self._bcp[self._bcp_cur]

    def _swap_bc_part(self):
        self._bcp_cur = 1 - self._bcp_cur
        self.work.bc_part = self._bcp[self._bcp_cur].data_ptr()
        self.work.bc_part_next = self._bcp[1 - self._bcp_cur].data_ptr()

    @property
    def gp(self):
        """Gradient records [N, (D+1) C] of the last call (unlimited gradient, then phi);
        with the fused stage only the rows of cells with a boundary facet are current."""
        return self._gpbuf[self._latest]

    def set_fused(self, on: bool):
        """Select the tile-fused stage (one launch per RHS; needs the mesh's tile tables)
        or the three-kernel pipeline.  Either may follow the other: both carry the
        boundary cells' records in the latest buffer."""
        self.fused = bool(on) and self.mesh.tiles is not None
        self.mesh_struct = self.mesh.struct if self.fused else self.mesh.struct3

    def _before_grad(self):
        """Point work.gp / gp_prev at the buffers of the next gradient launch."""
        self._last_has_prev = self.has_prev
        if self.fused:
            self.work.gp = self._gpbuf[1 - self._latest].data_ptr()
            self.work.gp_prev = self._gpbuf[self._latest].data_ptr()
        else:
            self.work.gp = self._gpbuf[self._latest].data_ptr()
            self.work.gp_prev = None

    def _after_grad(self):
        self._last_fused = self.fused
        if self.fused:
            self._latest = 1 - self._latest

    # The reference keeps the previous call's limited gradients in ``cell_grads``.
    @property
    def cell_grads(self):
        if not self.has_prev:
            return None
        D, C = self.dim, self.n_comp
        g = self._full_records().view(self.n_cells, D + 1, C)
        return g[:, :D, :] * g[:, D:D + 1, :]

    def _full_records(self):
        """Every cell's gradient record of the last call: the record buffer after the
        three-kernel pipeline; after the fused stage (which stores boundary cells'
        records only) recomputed from the last call's primitives and the records it
        read."""
        if not self._last_fused:
            return self.gp
        if self.last_prim is None:
            raise RuntimeError("cell_grads needs the primitives of the last evaluation (pass prim to stage())")
        if self._gfull is None:
            self._gfull = torch.empty_like(self._gpbuf[0])
        w = nat.FvmWork(ubc=self.ubc.data_ptr(), gp=self._gfull.data_ptr(), bc_part=self.bc_part.data_ptr(),
                        red=self.red.data_ptr(), fflux=self.fflux.data_ptr(),
                        gp_prev=self._gpbuf[1 - self._latest].data_ptr(),
                        bc_part_next=self._bcp[1 - self._bcp_cur].data_ptr())
        rc = nat.lib().fvm_grad_f32(self.mesh_struct, self.phys, w, nat.ptr(self.last_prim.contiguous()),
                                    int(self._last_has_prev), nat.stream_ptr(self.device))
        nat.check(rc, "fvm_grad_f32")
        self.bc_ready_ptr = None     # the pre-pass overwrote the boundary sums
        return self._gfull

    def reset_state(self):
        """Forget the previous call's gradients (the reference's ``cell_grads = None``)."""
        self.has_prev = False
        self.bc_ready_ptr = None

    def rhs(self, prim: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        nat.require_device(prim, "primatives")
        if prim.shape != (self.n_cells, self.n_comp) or prim.dtype != torch.float32:
            raise ValueError(f"primatives must be float32 [{self.n_cells}, {self.n_comp}]")
        prim = prim.contiguous()
        dudt = out if out is not None else torch.empty_like(prim)
        self._before_grad()
        rc = nat.lib().fvm_rhs_f32(self.mesh_struct, self.phys, self.work, nat.ptr(prim), nat.ptr(dudt),
                                   int(self.has_prev), nat.stream_ptr(self.device))
        nat.check(rc, "fvm_rhs_f32")
        self._after_grad()
        self.last_prim = prim
        self.has_prev = True
        self.bc_ready_ptr = None
        return dudt

    def stage(self, st: "nat.FvmStage", parts: int | None = None, prim: torch.Tensor | None = None):
        """One fused RK stage; the boundary means are reused when the previous stage's
        epilogue computed them for exactly this primitive buffer.  ``prim`` (the tensor
        behind st.prim) is remembered for saving."""
        self.last_prim = prim
        st.has_prev = int(self.has_prev)
        st.bc_ready = int(self.bc_ready_ptr is not None and self.bc_ready_ptr == st.prim)
        grad = parts is None or bool(parts & (2 | 16))
        if grad:
            self._before_grad()
        if parts is None:
            rc = nat.lib().fvm_rk_stage_f32(self.mesh_struct, self.phys, self.work, st, nat.stream_ptr(self.device))
        else:
            if not self.fused:
                parts &= ~16
            rc = nat.lib().fvm_rk_stage_part_f32(self.mesh_struct, self.phys, self.work, st, parts,
                                                 nat.stream_ptr(self.device))
        nat.check(rc, "fvm_rk_stage_f32")
        if grad:
            self._after_grad()
            self.has_prev = True
        if parts is None or parts & (8 | 16):
            # the epilogue wrote the next stage's partials into bc_part_next
            self._swap_bc_part()
            self.bc_ready_ptr = st.prim_next

    def invalidate_boundary_means(self):
        """Call after writing primitives the stage epilogue did not produce."""
        self.bc_ready_ptr = None

    def clear_temp(self):
        """The reference frees init-only tensors here (facet_process.py:290-295); the
        device layout holds none."""
        return None
