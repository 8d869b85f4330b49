"""Pin the CPU oracle against golden vectors written from the reference itself.

The fixtures (tests/golden/ref_*.npz) hold the reference's own outputs for small seeded
meshes; see tests/golden/make_golden.py.  Index/ordering arrays must match exactly;
floating-point arrays are compared to fp32 round-off (the restatement reproduces the
reference's op sequence, so in practice they match bit for bit).
"""
import numpy as np
import pytest
import torch

from oracle import fvm_oracle as O
from tests.golden_cases import load, mesh_tuple, phys_of

CASES = ["ref_2d.npz", "ref_3d.npz"]
RTOL = 1e-6


def _close(a, b, rtol=RTOL):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    assert a.shape == b.shape
    scale = max(np.max(np.abs(b)), 1e-30)
    assert np.max(np.abs(a - b)) <= rtol * scale


def _exact(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape
    assert np.array_equal(a.astype(np.int64), b.astype(np.int64))


@pytest.fixture(scope="module", params=CASES)
def case(request):
    d = load(request.param)
    mesh, fvm = O.from_mesh_tuple(*mesh_tuple(d), phys_of(d))
    return d, mesh, fvm


def test_mesh_topology_exact(case):
    d, mesh, fvm = case
    _exact(mesh.cell_to_facet, d["cell_to_facet"])
    _exact(mesh.cell_facet_signs, d["cell_facet_signs"])
    _exact(mesh.facet_to_cell_main, d["facet_to_cell_main"])
    _exact(mesh.facet_to_cell_bc, d["facet_to_cell_bc"])
    _exact(mesh.combined_neigh, d["combined_neigh"])
    _exact(fvm.slots, d["slots"])
    _exact(fvm.bc_facet_side, d["bc_facet_side"])
    _exact(fvm.bc_locations, d["bc_locations"])


def test_mesh_geometry(case):
    d, mesh, fvm = case
    for k, v in [("normals", mesh.normals), ("midpoints", mesh.midpoints), ("volumes", mesh.volumes),
                 ("centroids", mesh.centroids), ("cent_to_facet_disp", mesh.cent_to_facet_disp),
                 ("cell_disps", mesh.cell_disps), ("facet_dist_bc", mesh.facet_dist_bc)]:
        _close(v, d[k])
    _close(fvm.X_orthog.squeeze(-1), d["X_orthog"])
    _close(fvm.normals_hat.squeeze(-1), d["normals_hat"])
    _close(fvm.eps_p.reshape(-1), d["eps_p"])


@pytest.mark.parametrize("name", ["flux_mat", "G_mats", "A_face_grad", "A_bc", "A_corr"])
def test_operators(case, name):
    d, mesh, fvm = case
    op = {"flux_mat": fvm.flux_mat, "G_mats": fvm.G, "A_face_grad": fvm.A_face_grad,
          "A_bc": fvm.A_bc, "A_corr": fvm.A_corr}[name]
    crow, col, val = op.csr()
    _exact(crow, d[name + "_crow"])
    _exact(col, d[name + "_col"])
    _close(val, d[name + "_val"])
    ev, ec = O.csr_to_ell(crow, col, val)
    _exact(ec, d[name + "_ell_col"])
    _close(ev, d[name + "_ell_val"])


def test_bias_vectors(case):
    d, mesh, fvm = case
    _close(fvm.b_face_grad, d["b_face_grad"])
    b = fvm.b_bc.numpy()
    defined = ~np.isnan(b)            # other rows come from torch.empty in the reference
    _close(b[defined], d["b_bc"][defined])


def test_rhs_two_calls(case):
    d, mesh, fvm = case
    fvm.cell_grads = None
    for call in range(2):
        prim = torch.from_numpy(d[f"rhs{call}_prim"])
        _close(fvm.set_face_values(prim, fvm.cell_grads), d[f"rhs{call}_ubc"])
        dudt = fvm.forward(prim)
        _close(fvm.cell_grads, d[f"rhs{call}_grads"])
        _close(fvm.grad_V, d[f"rhs{call}_grad_V"])
        _close(fvm.grad_T_n, d[f"rhs{call}_grad_T_n"])
        _close(fvm.fluxes, d[f"rhs{call}_fluxes"])
        _close(dudt, d[f"rhs{call}_dudt"])


@pytest.mark.parametrize("name,extra", [("Butcher", "RK4"), ("Butcher_adapt", "RK3_SSP4"), ("RK3_SSP4", "")])
@pytest.mark.parametrize("fname", CASES)
def test_integrator_steps(fname, name, extra):
    d = load(fname)
    mesh, fvm = O.from_mesh_tuple(*mesh_tuple(d), phys_of(d))
    st = O.OracleStepper(fvm, name, extra, dt=float(d["step_dt"]))
    U = torch.from_numpy(d["step_U0"])
    for i in range(10):
        U = st.step(U)
        _close(U, d[f"step_{name}{extra}_U"][i], rtol=1e-6)
        assert abs(float(st.dt) - float(d[f"step_{name}{extra}_dt"][i])) <= 1e-6 * float(st.dt)


def test_csr_to_ell_random():
    d = load("ref_csr_to_ell.npz")
    for i in range(3):
        crow = torch.from_numpy(d[f"csr{i}_crow"])
        col = torch.from_numpy(d[f"csr{i}_col"])
        val = torch.from_numpy(d[f"csr{i}_val"])
        v, c = O.csr_to_ell(crow, col, val)
        _exact(c, d[f"csr{i}_ell_col"])
        _close(v, d[f"csr{i}_ell_val"], rtol=0)
