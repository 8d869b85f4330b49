"""FP64 geometry on the GPU (csrc/geom.hip) against the host float64 computation of the
same mesh: floating values to fp64 round-off, index structures identical, and the
packed fp32 device layout identical."""
import numpy as np
import pytest

from fvm_solver_amd.config_fvm import ConfigChannel, ConfigChannel3D
from fvm_solver_amd.fvm_stepping.facet_process import host_layout
from fvm_solver_amd.mesh_gen.channel import channel_mesh_2d, channel_mesh_3d
from fvm_solver_amd.mesh_utils.fvm_mesh import mesh_from_tuple
from fvm_solver_amd.mesh_utils.mesh_store import channel_bc_tags

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,n", [(2, 20000), (3, 20000)])
def test_gpu_geometry_matches_host(dim, n):
    gen = channel_mesh_2d if dim == 2 else channel_mesh_3d
    pts, cells, (fi, fb), tags = gen(n, seed=3)
    mh = mesh_from_tuple(pts, cells, fi, fb, device="cpu")
    mg = mesh_from_tuple(pts, cells, fi, fb, device="cuda:0")
    for k in ("normals", "normals_hat", "midpoints", "volumes", "centroids", "cent_to_facet_disp", "G"):
        a, b = mg.geo64[k], mh.geo64[k]
        err = np.abs(a - b).max() / np.abs(b).max()
        print(f"[gpu-test]   {dim}-D {k}: max rel diff {err:.1e}")
        assert err < 1e-12, k
    for k in ("signs", "left", "right", "c2f", "combined", "is_bc"):
        assert np.array_equal(mg.geo64[k], mh.geo64[k]), k
    cfg = ConfigChannel(device="cpu") if dim == 2 else ConfigChannel3D(device="cpu")
    lh = host_layout(mh, dim + 2, channel_bc_tags(tags, dim), cfg)
    lg = host_layout(mg, dim + 2, channel_bc_tags(tags, dim), cfg)
    for k in ("info", "nbr_s", "c2f_s"):
        assert np.array_equal(lg[k], lh[k]), k
    for k in ("gval", "geo", "inv_vol", "bc_val", "bc_ga"):
        diff = np.abs(lg[k].astype(np.float64) - lh[k]).max() / max(np.abs(lh[k]).max(), 1e-30)
        assert diff < 1e-6, (k, diff)
