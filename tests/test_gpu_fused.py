"""The tile-fused stage (one launch per RHS), with every workgroup size the mesh allows
(256 to 512 threads per tile), against the three-kernel pipeline.

All three evaluate the same device functions on the same floats in the same order, so every
result must be bitwise identical: dU/dt of the reference's stateful forward over two
calls (the second applies the Neumann correction with the first call's gradients), the
carried gradients (cell_grads), RK4 and embedded-tableau trajectories, batches of
meshes, and graph replays.  Parity of either pipeline with the CPU oracle is covered by
test_gpu_parity.py, which runs the default (fused) pipeline.
"""
import numpy as np
import pytest
import torch

from fvm_solver_amd.fvm_stepping.facet_process import FacetCalc
from fvm_solver_amd.mesh_gen.channel import channel_mesh_2d, channel_mesh_3d
from tests.golden_cases import load, mesh_tuple
from tests.gpu_helpers import DEV, announce, initial_state, make_cfg, product_equation

pytestmark = pytest.mark.gpu


def _announce(msg):
    announce(msg)


@pytest.mark.parametrize("fname", ["ref_2d.npz", "ref_3d.npz"])
def test_rhs_two_calls_bitwise(fname):
    d = load(fname)
    pts, cells, fi, fb, tags = mesh_tuple(d)
    cfg = make_cfg(cells.shape[1] - 1)
    mesh, eq = product_equation(pts, cells, fi, fb, tags, cfg)
    dm = eq.E_props.mesh
    assert dm.tiles is not None, dm.tiles_reason
    _announce(f"fused vs three-kernel rhs, {dm.N} cells, lds {dm.tiles['lds_bytes']} B")
    out = {}
    opts = dm.fused_thread_options()
    assert 256 in opts and 512 in opts, opts
    for fused in (False, *opts):
        if fused:
            dm.set_fused_threads(fused)
        E = FacetCalc.from_device_mesh(dm, cfg, DEV, fused=bool(fused))
        assert E.fused == bool(fused)
        res = []
        for call in (0, 1):
            prim = torch.from_numpy(d[f"rhs{call}_prim"]).to(DEV)
            res.append(E.rhs(prim).clone())
            res.append(E.cell_grads.clone())
            res.append(E.ubc.clone())
        torch.cuda.synchronize()
        out[fused] = res
    for nt in opts:
        for k, (a, b) in enumerate(zip(out[False], out[nt])):
            assert torch.equal(a, b), (nt, k, float((a - b).abs().max()))


def _channel_eq(n, solver, extra, seed=4, dim=2):
    gen = channel_mesh_2d if dim == 2 else channel_mesh_3d
    pts, cells, (fi, fb), tags = gen(n, seed=seed)
    cfg = make_cfg(dim, solver_name=solver, solver_extra=extra)
    c = pts[cells].mean(1)
    _, U0 = initial_state(c, cfg, dim)
    mesh, eq = product_equation(pts, cells, fi, fb, tags, cfg, us_init=U0)
    return eq


@pytest.mark.parametrize("dim,n", [(2, 20000), (3, 20000), (3, 300000)])
@pytest.mark.parametrize("solver,extra", [("Butcher", "RK4"), ("Butcher_adapt", "RK3_SSP4")])
def test_trajectories_bitwise(solver, extra, dim, n):
    states = []
    opts = [256, 320, 384, 512]
    for fused in (False, *opts):
        eq = _channel_eq(n, solver, extra, dim=dim)
        if fused and fused not in eq.E_props.mesh.fused_thread_options():
            continue
        if fused:
            assert eq.E_props.mesh.tiles is not None, eq.E_props.mesh.tiles_reason
            eq.E_props.mesh.set_fused_threads(fused)
        eq.E_props.set_fused(bool(fused))
        assert eq.E_props.fused == bool(fused)
        ts = eq.t_solver
        ts.set_dt(float(eq.cfl_dt(eq.cells.get_values()[0], 0.5)))
        dts = []
        for _ in range(12):
            ts._solve_step(None)
            dts.append(float(ts.dt))
        torch.cuda.synchronize()
        states.append((eq.cells.state.clone(), dts, eq.E_props.cell_grads.clone()))
    _announce(f"{dim}-D {n} cells {solver} {extra}: 12 steps fused ({len(states) - 1} workgroup sizes) vs "
              f"three-kernel, dt {states[1][1][-1]:.4e}")
    assert len(states) >= 3
    for st in states[1:]:
        assert torch.isfinite(st[0]).all()
        assert torch.equal(states[0][0], st[0])
        assert states[0][1] == st[1]
        assert torch.equal(states[0][2], st[2])


def test_switching_pipelines_mid_run():
    """The two pipelines may follow each other: both carry the boundary cells'
    records, so alternating them gives the same trajectory as either alone."""
    ref = _channel_eq(8000, "Butcher", "RK4")
    alt = _channel_eq(8000, "Butcher", "RK4")
    for eq in (ref, alt):
        eq.t_solver.set_dt(float(eq.cfl_dt(eq.cells.get_values()[0], 0.5)))
    for k in range(6):
        alt.E_props.mesh.set_fused_threads(512 if k % 4 == 2 else 256)
        alt.E_props.set_fused(k % 2 == 0)
        ref.t_solver._solve_step(None)
        alt.t_solver._solve_step(None)
    torch.cuda.synchronize()
    assert torch.equal(ref.cells.state, alt.cells.state)


def test_batch_and_graph_bitwise():
    from fvm_solver_amd.batch import BatchedSolver, random_channel_batch
    problems, params = random_channel_batch(5, seed=7, n_min=2000, n_max=12000, device=DEV)
    res = []
    for fused in (False, 256, 320, 384, 512):
        bs = BatchedSolver(problems, device=DEV)
        if fused and fused not in bs.E.mesh.fused_thread_options():
            continue
        if fused:
            bs.E.mesh.set_fused_threads(fused)
        bs.E.set_fused(bool(fused))
        bs.set_dt(0.5 * bs.cfl_dt(0.5))
        for _ in range(3):
            bs.step()
        g = bs.capture_graph(2)
        snaps = []
        for _ in range(3):
            g.replay()
            torch.cuda.synchronize()
            snaps.append(bs.state.clone())
        res.append(snaps)
    _announce(f"batch of {[p['cells'] for p in params]} cells: fused ({len(res) - 1} workgroup sizes) vs "
              "three-kernel, graph replays")
    for other in res[1:]:
        for a, b in zip(res[0], other):
            assert torch.isfinite(b).all()
            assert torch.equal(a, b)
