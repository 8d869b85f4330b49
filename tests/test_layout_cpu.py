"""Host-side checks of the product's mesh preprocessing and device layout (no GPU).

The product computes geometry in float64 (the reference in float32), so index and
ordering arrays must match the oracle exactly while values agree to float32 round-off
amplified by the LSQ conditioning.  The operators the reference assembles (flux_mat,
G_mats, A_face_grad + b, A_bc + b_bc, A_corr) are rebuilt from the packed layout and
compared entry by entry with the oracle's.
"""
import numpy as np
import pytest
import torch

from fvm_solver_amd.config_fvm import ConfigChannel, ConfigChannel3D
from fvm_solver_amd.fvm_stepping.facet_process import DeviceMesh
from fvm_solver_amd.mesh_utils.fvm_mesh import mesh_from_tuple
from fvm_solver_amd.mesh_utils.mesh_store import channel_bc_tags
from oracle import fvm_oracle as O
from tests.golden_cases import load, mesh_tuple, phys_of

VAL_RTOL = 2e-5


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30)


@pytest.fixture(scope="module", params=["ref_2d.npz", "ref_3d.npz"])
def case(request):
    d = load(request.param)
    pts, cells, fi, fb, tags = mesh_tuple(d)
    dim = cells.shape[1] - 1
    cfg = ConfigChannel() if dim == 2 else ConfigChannel3D()
    mesh = mesh_from_tuple(pts, cells, fi, fb, device="cpu")
    dm = DeviceMesh(mesh, dim + 2, channel_bc_tags(tags, dim), cfg, "cpu")
    omesh, ofvm = O.from_mesh_tuple(pts, cells, fi, fb, tags, phys_of(d))
    return d, mesh, dm, omesh, ofvm


def test_topology_matches_reference(case):
    d, mesh, dm, om, of = case
    for name in ("cell_to_facet", "cell_facet_signs", "facet_to_cell_main", "facet_to_cell_bc"):
        assert np.array_equal(getattr(mesh, name).numpy(), d[name]), name
    assert np.array_equal(mesh.cell_grad_stuff[3].numpy(), d["combined_neigh"])
    assert np.array_equal(dm.host["side"], d["bc_facet_side"])


def test_geometry_close(case):
    d, mesh, dm, om, of = case
    for name in ("normals", "midpoints", "volumes", "centroids"):
        assert _rel(getattr(mesh, name), d[name]) < 1e-6, name
    # differences of O(1) coordinates: the fp32 reference carries ~1 ulp of the
    # coordinates (not of the difference), so compare absolutely
    cd, dist, G, comb = mesh.cell_grad_stuff
    for mine, ref in ((mesh.cent_to_facet_disp, d["cent_to_facet_disp"]), (cd, d["cell_disps"]),
                      (dist, d["facet_dist_bc"])):
        assert np.max(np.abs(np.asarray(mine) - ref)) < 5e-7


def _csr_dense(op):
    crow, col, val = op.csr()
    M = crow.numel() - 1
    rows = torch.repeat_interleave(torch.arange(M), crow[1:] - crow[:-1])
    return rows.numpy(), col.numpy(), val.numpy()


def test_flux_matrix_layout(case):
    d, mesh, dm, om, of = case
    h = dm.host
    N, NL = mesh.n_cells, mesh.dim + 1
    rows, cols, vals = _csr_dense(of.flux_mat)
    # ELL row order == ascending facet order, same facets, same signs, |val| = 1/vol
    assert np.array_equal(cols.reshape(N, NL), h["c2f_s"])
    sgn = np.where(h["slot_s"] == 1, 1.0, -1.0)
    mine = sgn * h["inv_vol"][:, None]
    assert _rel(mine, vals.reshape(N, NL)) < VAL_RTOL


def test_gradient_layout(case):
    d, mesh, dm, om, of = case
    h = dm.host
    N, D, NL = mesh.n_cells, mesh.dim, mesh.dim + 1
    rows, cols, vals = _csr_dense(of.G)
    cols = cols.reshape(N, D, NL + 1)
    vals = vals.reshape(N, D, NL + 1)
    # gval holds the oracle's G row in slot order: the column of each slot's neighbour
    # (or boundary value), then the cell's own column
    col_mine = np.concatenate([h["nbr_s"], np.arange(N)[:, None]], 1)
    for i in range(D):
        # every oracle column appears once among the cell's slots + itself
        assert np.array_equal(np.sort(cols[:, i, :], 1), np.sort(col_mine, 1))
    pos = np.argsort(cols[:, 0, :], 1, kind="stable")
    want = np.take_along_axis(vals, pos[:, None, :], 2)             # oracle values by column id
    order = np.argsort(np.argsort(col_mine, 1, kind="stable"), 1)    # rank of each slot's column
    want = np.take_along_axis(want, order[:, None, :], 2)           # ... in slot order
    scale = np.abs(vals).max(axis=2, keepdims=True)
    assert np.max(np.abs(h["gval"] - want) / scale) < VAL_RTOL


def test_neighbour_back_pointers(case):
    d, mesh, dm, om, of = case
    h = dm.host
    N = mesh.n_cells
    nb = h["nbr_s"]
    inter = nb < N
    c, k = np.nonzero(inter)
    assert np.array_equal(h["c2f_s"][nb[c, k], h["nbk"][c, k]], h["c2f_s"][c, k])


def test_boundary_layout(case):
    d, mesh, dm, om, of = case
    h = dm.host
    N, D, C, Fb = mesh.n_cells, mesh.dim, mesh.dim + 2, mesh.n_bc_facet
    # A_bc + b_bc (defined rows), A_face_grad boundary rows + b, A_corr
    rows, cols, vals = _csr_dense(of.A_bc)
    neum = np.zeros(Fb * C, bool)
    neum[rows] = True
    flags = ((h["flags"][:, None] >> np.arange(C)) & 1).astype(bool).reshape(-1)
    assert np.array_equal(neum, flags)
    bb = of.b_bc.numpy()
    defined = ~np.isnan(bb)
    assert _rel(h["bc_val"].reshape(-1)[defined], bb[defined]) < VAL_RTOL
    fidx = np.flatnonzero(mesh.bc_facet_mask.numpy())
    rows, cols, vals = _csr_dense(of.A_face_grad)
    A = np.zeros((mesh.n_facets * C,))
    A_cells = np.full((mesh.n_facets * C,), -1)
    bsel = np.isin(rows // C, fidx)
    A[rows[bsel]] = vals[bsel]
    A_cells[rows[bsel]] = cols[bsel] // C
    ga = A[(fidx[:, None] * C + np.arange(C)).reshape(-1)].reshape(Fb, C)
    assert _rel(h["bc_ga"], ga) < VAL_RTOL
    gb = of.b_face_grad.numpy()[(fidx[:, None] * C + np.arange(C)).reshape(-1)].reshape(Fb, C)
    assert _rel(h["bc_gb"], gb) < VAL_RTOL
    rows, cols, vals = _csr_dense(of.A_corr)
    b_of_row = rows // C
    l_ref = np.zeros((Fb, D))
    l_ref[b_of_row, (cols // C) % D] = vals
    assert _rel(h["bc_l"], l_ref) < VAL_RTOL
    g = of.groups[0]
    idx = g["bc_idx"].numpy()
    ch = h["bc_char"][idx]
    assert np.all(h["group"][idx] == 0) and np.sum(h["group"] >= 0) == len(idx)
    assert _rel(ch[:, :D], g["normals"]) < VAL_RTOL
    assert _rel(ch[:, D], g["rho_inf"]) == 0
    assert _rel(ch[:, D + 1], g["v_n_inf"]) < VAL_RTOL
    assert _rel(ch[:, D + 2], g["p_inf"]) == 0
    assert np.max(np.abs(ch[:, D + 3:] - g["v_t_inf"].numpy())) < 1e-5


def test_facet_geometry_layout(case):
    d, mesh, dm, om, of = case
    h = dm.host
    D = mesh.dim
    geo = h["geo"]
    assert _rel(geo[:, :D], of.normals) < VAL_RTOL
    assert _rel(geo[:, D:2 * D], of.normals_hat.squeeze(-1)) < VAL_RTOL
    assert _rel(geo[:, 2 * D], of.facet_size.squeeze(-1)) < VAL_RTOL
    inter = ~mesh.bc_facet_mask.numpy()
    assert _rel(geo[inter, 2 * D + 1], 1.0 / of.cell_dist_proj[torch.from_numpy(inter)].numpy()) < VAL_RTOL
    X = of.X_orthog.squeeze(-1).numpy()
    assert np.max(np.abs(geo[:, 2 * D + 2:3 * D + 2] - X)) < 1e-4 * np.abs(X).max()
    assert _rel(h["eps"], of.eps_p.reshape(-1)) < VAL_RTOL


def test_facet_record_side_displacements(case):
    d, mesh, dm, om, of = case
    D = mesh.dim
    geo = dm.host["geo"]
    disp = d["cent_to_facet_disp"]
    c2f = d["cell_to_facet"]
    slots = d["slots"].reshape(c2f.shape)
    for s_ in (0, 1):
        c, k = np.nonzero(slots == s_)
        got = geo[c2f[c, k], 3 * D + 2 + s_ * D:3 * D + 2 + (s_ + 1) * D]
        assert np.max(np.abs(got - disp[c, k])) < 5e-7


def test_facet_sides_and_device_order(case):
    """fac_cell names each facet's slot-0 / slot-1 cell as the reference's facet_to_cell
    tables and facet signs do (outer side of a boundary facet: N + boundary index), and
    the device facet order is a permutation sorted by lowest adjacent cell."""
    d, mesh, dm, om, of = case
    h = dm.host
    N, F = dm.N, dm.F
    fc = h["fac_cell"]
    c2f = d["cell_to_facet"]
    slots = d["slots"].reshape(c2f.shape)
    for s_ in (0, 1):
        c, k = np.nonzero(slots == s_)
        assert np.array_equal(fc[s_, c2f[c, k]], c)
    bmask = mesh.bc_facet_mask.numpy()
    fb = np.flatnonzero(bmask)
    outer = np.where(fc[0, fb] >= N, fc[0, fb], fc[1, fb])
    assert np.array_equal(outer, N + np.arange(len(fb)))
    inter = np.flatnonzero(~bmask)
    assert (fc[:, inter] < N).all() and (fc[0, inter] != fc[1, inter]).all()
    perm, pos = h["fperm"], h["fpos"]
    assert np.array_equal(np.sort(perm), np.arange(F)) and np.array_equal(pos[perm], np.arange(F))
    # device order: by the 256-cell run of the lowest adjacent cell, boundary facets of a
    # run last, then by lowest cell
    lo, bnd = fc.min(0)[perm], (fc.max(0) >= N)[perm]
    key = (lo // 256) * 2 * (N + 1) + bnd * (N + 1) + lo
    assert (np.diff(key) >= 0).all()


def test_batch_layout_offsets_and_group_tables():
    from fvm_solver_amd.config_fvm import channel_config
    from fvm_solver_amd.mesh_gen.channel import ChannelSpec, channel_mesh_2d
    items, meshes = [], []
    for i, (n, mach, r) in enumerate(((600, 0.2, 0.08), (300, 0.5, 0.12), (900, 0.35, 0.1))):
        pts, cells, (fi, fb), tags = channel_mesh_2d(n, seed=20 + i, spec=ChannelSpec(centre=(0.5, 0.5), radius=r))
        mesh = mesh_from_tuple(pts, cells, fi, fb, device="cpu")
        meshes.append((mesh, tags))
        items.append((mesh, 4, channel_bc_tags(tags, 2), channel_config(mach, 2 * r, device="cpu")))
    dm = DeviceMesh.batch(items, "cpu")
    singles = [DeviceMesh(*it, "cpu") for it in items]
    h = dm.host
    co, fo, bo = dm.cell_offsets, dm.facet_offsets, dm.bc_offsets
    for i, sd in enumerate(singles):
        a, b = co[i], co[i + 1]
        sh = sd.host
        assert np.array_equal(h["c2f_s"][a:b], sh["c2f_s"] + fo[i])
        nb = sh["nbr_s"]
        exp = np.where(nb < sd.N, nb + co[i], nb - sd.N + bo[i] + dm.N)
        assert np.array_equal(h["nbr_s"][a:b], exp)
        assert np.array_equal(h["info"][a:b], sh["info"])
        assert np.array_equal(h["geo"][fo[i]:fo[i + 1]], sh["geo"])
        assert np.array_equal(h["cell_mesh"][a:b], np.full(b - a, i))
        assert np.allclose(dm.mesh_phys[i, :3], [items[i][3].viscosity, items[i][3].visc_bulk,
                                                  items[i][3].thermal_cond])
    # every group facet's tile is a slot whose group window contains the group, and the
    # group's slot range is exactly the slots whose tiles touch its mesh
    g = h["group"]
    tiles = h["bcell"] // 256
    for b in np.flatnonzero(g >= 0):
        sl = h["blk_slot"][tiles[b]]
        assert sl >= 0 and h["slot_g0"][sl] <= g[b] < h["slot_g0"][sl] + h["slot_ng"][sl]
        assert h["group_s0"][g[b]] <= sl < h["group_s1"][g[b]]
    assert np.all(h["group_count"] == np.bincount(g[g >= 0]))
