"""A mesh in arbitrary cell order, as gmsh writes them (run_fvm.py:14-34): a channel mesh
with its cells shuffled, written as a .msh file and read back, goes through the drop-in
FVMEquation.  Its own order defeats the tile tables, so the equation renumbers the cells
on the device (Morton order) and runs the tile-fused stage; forward() takes and returns
the mesh's order, the solver's state maps back with state_in_mesh_order().  Both are
checked against the fp32 oracle on the shuffled mesh itself: the RHS on two successive
calls and 20 RK4 steps (<= 1e-4 relative L2 per conserved variable), and the same run
with reordering switched off (three-kernel pipeline) to the same bar.
"""
import numpy as np
import pytest
import torch

from fvm_solver_amd.mesh_gen.channel import channel_mesh_2d
from fvm_solver_amd.mesh_gen.gmsh_io import read_msh, write_msh
from oracle import fvm_oracle as O
from tests.gpu_helpers import DEV, announce, initial_state, make_cfg, oracle_phys, product_equation, rel_l2_cols

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def shuffled_msh(tmp_path_factory):
    pts, cells, (fi, fb), tags = channel_mesh_2d(20000, seed=8)
    perm = np.random.default_rng(4).permutation(len(cells))
    path = str(tmp_path_factory.mktemp("msh") / "channel_shuffled.msh")
    write_msh(path, pts, cells[perm], fb, tags)
    return read_msh(path)


@pytest.mark.parametrize("reorder", ["auto", "never"])
def test_shuffled_gmsh_mesh_against_oracle(shuffled_msh, reorder):
    pts, cells, (fi, fb), tags = shuffled_msh
    cfg = make_cfg(2, solver_name="Butcher", solver_extra="RK4", reorder=reorder)
    prim0, U0 = initial_state(pts[cells].mean(1), cfg, 2)
    mesh, eq = product_equation(pts, cells, fi, fb, tags, cfg, us_init=U0)
    if reorder == "auto":
        assert eq.cell_perm is not None and eq.E_props.fused, eq.E_props.mesh.tiles_reason
    else:
        assert eq.cell_perm is None and not eq.E_props.fused
    announce(f"shuffled .msh mesh, {mesh.n_cells} cells, reorder={reorder}, fused={eq.E_props.fused}")
    _, o32 = O.from_mesh_tuple(pts, cells, fi, fb, tags, oracle_phys(cfg))
    # the stateful forward in the mesh's order, two calls
    for call in range(2):
        p = prim0 * (1.0 + 0.01 * call)
        mine = eq.forward(p.to(DEV)).cpu().numpy()
        ref = o32.forward(p).numpy()
        err = rel_l2_cols(mine, ref)
        announce(f"  rhs call {call}: rel L2 vs oracle {err}")
        assert np.all(err <= 1e-4), err
    # 20 RK4 steps from the initial state (fresh carried gradients on both sides)
    eq.E_props.reset_state()
    o32.cell_grads = None
    dt = float(eq.cfl_dt(prim0.to(DEV), 0.5))
    eq.t_solver.set_dt(dt)
    st = O.OracleStepper(o32, "Butcher", "RK4", dt=dt)
    U = U0.clone()
    for _ in range(20):
        eq.t_solver._solve_step(None)
        U = st.step(U)
    err = rel_l2_cols(eq.state_in_mesh_order().cpu().numpy(), U.numpy())
    announce(f"  20 RK4 steps: rel L2 vs oracle {err}")
    assert np.all(err <= 1e-4), err
    tot = eq.conservation_totals().cpu().numpy()
    ref = (mesh.geo64["volumes"][:, None] * eq.state_in_mesh_order().cpu().double().numpy()).sum(0)
    assert np.allclose(tot, ref, rtol=1e-12)
