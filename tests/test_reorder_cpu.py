"""Cell reordering (no GPU): permutation validity, bandwidth reduction on a shuffled mesh,
and that the packed operators of the reordered mesh are those of the original mesh with
renumbered cells (index structures bit-exact after un-permuting)."""
import numpy as np
import pytest

from fvm_solver_amd.config_fvm import ConfigChannel
from fvm_solver_amd.fvm_stepping.facet_process import DeviceMesh
from fvm_solver_amd.mesh_gen.channel import channel_mesh_2d
from fvm_solver_amd.mesh_utils.fvm_mesh import mesh_from_tuple
from fvm_solver_amd.mesh_utils.mesh_store import channel_bc_tags
from fvm_solver_amd.mesh_utils.reorder import bandwidth, mean_index_distance, reorder_cells


@pytest.fixture(scope="module")
def shuffled():
    pts, cells, (fi, fb), tags = channel_mesh_2d(3000, seed=4)
    rng = np.random.default_rng(0)
    sh = rng.permutation(len(cells))
    return pts, cells[sh], fi, fb, tags


@pytest.mark.parametrize("method", ["rcm", "morton"])
def test_reorder_is_a_permutation_and_narrows_bandwidth(shuffled, method):
    pts, cells, fi, fb, tags = shuffled
    new, perm = reorder_cells(pts, cells, method)
    assert np.array_equal(np.sort(perm), np.arange(len(cells)))
    assert np.array_equal(new, cells[perm])
    m0, m1 = mean_index_distance(cells), mean_index_distance(new)
    assert m1 < 0.05 * m0, (m0, m1)
    if method == "rcm":
        assert bandwidth(new) < 0.1 * bandwidth(cells)


def test_operators_match_after_unpermuting(shuffled):
    pts, cells, fi, fb, tags = shuffled
    new, perm = reorder_cells(pts, cells, "rcm")
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    cfg = ConfigChannel(device="cpu")
    m0 = mesh_from_tuple(pts, cells, fi, fb, device="cpu")
    m1 = mesh_from_tuple(pts, new, fi, fb, device="cpu")
    d0 = DeviceMesh(m0, 4, channel_bc_tags(tags, 2), cfg, "cpu").host
    d1 = DeviceMesh(m1, 4, channel_bc_tags(tags, 2), cfg, "cpu").host
    N = len(cells)
    # facet lists per cell are facet ids (unchanged by a cell renumbering)
    assert np.array_equal(d1["c2f_s"], d0["c2f_s"][perm])
    # neighbour lists: renumbered cells, boundary entries unchanged
    nb0 = d0["nbr_s"][perm]
    exp = np.where(nb0 < N, inv[np.minimum(nb0, N - 1)], nb0)
    assert np.array_equal(d1["nbr_s"], exp)
    assert np.array_equal(d1["slot_s"], d0["slot_s"][perm])
    assert np.allclose(d1["inv_vol"], d0["inv_vol"][perm], rtol=0, atol=0)
    assert np.array_equal(d1["geo"], d0["geo"])
