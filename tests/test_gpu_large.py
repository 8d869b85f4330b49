"""Full-size GPU runs of the benchmark configurations (run_fvm.py:115-150 at the
benchmark sizes), and a batch of meshes against the oracle mesh by mesh.

  * 2-D channel, ~1M cells: the tile-fused stage (every workgroup size of 256 to 512
    threads per tile the tables allow) against the three-kernel pipeline, bitwise, over RK4 and AB3 steps (3.9k tiles: far more than can be resident, so every
    ordering between tiles of one launch is exercised); 10 RK4 steps against the fp32
    oracle (<= 1e-4 relative L2 per conserved variable where the fp32 oracle is within
    1e-5 of the fp64 one; rho v, which starts at zero, no further from fp64 than the fp32
    oracle); 2000 RK4 steps at CFL 0.5 and
    2000 AB3 steps at CFL 0.2 checked for finiteness and conservation: the fp64 change
    of sum_c vol_c U_c against the trapezoid integral of the boundary flux, to 1e-6 of
    the gross flux sum_f |F_f| integrated over the run (measured: <= 1.5e-8).
  * Batch: 6 channel meshes of different Mach number, obstacle and size (one above 100k
    cells) in one block-diagonal layout, 20 RK4 steps, each mesh against its own fp32
    oracle run (<= 1e-4 per variable).
  * The 64-mesh dataset batch (5k-200k cells each): 200 RK4 steps, finite, per-mesh
    conservation as above, sampled every second step, to 3e-5 of the gross flux
    (measured: <= 2.3e-6).
  * 3-D channel, ~4M tetrahedra: 5 RK4 steps against the fp32 and fp64 oracles, 500
    steps with the conservation ledger (measured: <= 4.8e-8 of the gross flux).
"""
import numpy as np
import pytest
import torch

from fvm_solver_amd.mesh_gen.channel import channel_mesh_2d
from oracle import fvm_oracle as O
from tests.gpu_helpers import (DEV, FluxLedger, announce, initial_state, make_cfg, oracle_phys, product_equation,
                               rel_l2_cols)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def channel_1m():
    announce("generating the 1M-cell channel mesh")
    return channel_mesh_2d(1_000_000, seed=1)


def _eq_1m(mesh_t, solver, extra, cfl):
    pts, cells, (fi, fb), tags = mesh_t
    cfg = make_cfg(2, solver_name=solver, solver_extra=extra)
    _, U0 = initial_state(pts[cells].mean(1), cfg, 2)
    mesh, eq = product_equation(pts, cells, fi, fb, tags, cfg, us_init=U0)
    dt = float(eq.cfl_dt(eq.cells.get_values()[0], cfl))
    eq.t_solver.set_dt(dt)
    return cfg, mesh, eq, U0, dt


@pytest.mark.timeout(600)
@pytest.mark.parametrize("solver,extra,cfl", [("Butcher", "RK4", 0.5), ("AB3", "", 0.2)])
def test_fused_vs_three_kernels_1m_bitwise(channel_1m, solver, extra, cfl):
    states = []
    for fused in (False, 256, 320, 384, 512):
        cfg, mesh, eq, U0, dt = _eq_1m(channel_1m, solver, extra, cfl)
        tl = eq.E_props.mesh.tiles
        assert tl is not None, eq.E_props.mesh.tiles_reason
        if fused and fused not in eq.E_props.mesh.fused_thread_options():
            continue
        if fused:
            eq.E_props.mesh.set_fused_threads(fused)
        eq.E_props.set_fused(bool(fused))
        for _ in range(3):
            eq.t_solver._solve_step(None)
        torch.cuda.synchronize()
        states.append(eq.cells.state.clone())
    announce(f"{solver}{extra} 1M cells ({tl['n_tiles']} tiles): 3 steps fused ({len(states) - 1} workgroup sizes) "
             "vs three kernels")
    for st in states[1:]:
        assert torch.isfinite(st).all()
        assert torch.equal(states[0], st)


@pytest.mark.timeout(900)
def test_rk4_1m_against_oracle(channel_1m):
    cfg, mesh, eq, U0, dt = _eq_1m(channel_1m, "Butcher", "RK4", 0.5)
    pts, cells, (fi, fb), tags = channel_1m
    announce(f"RK4 1M cells ({mesh.n_cells}), dt {dt:.4e}: 10 steps vs fp32 oracle")
    for _ in range(10):
        eq.t_solver._solve_step(None)
    mine = eq.cells.state.cpu().numpy()
    torch.set_num_threads(max(1, min(16, torch.get_num_threads())))
    res = {}
    for name, dtype in (("fp32", torch.float32), ("fp64", torch.float64)):
        _, ofvm = O.from_mesh_tuple(pts, cells, fi, fb, tags, oracle_phys(cfg), dtype=dtype)
        st = O.OracleStepper(ofvm, "Butcher", "RK4", dt=dt)
        U = U0.to(dtype)
        for i in range(10):
            U = st.step(U)
        res[name] = U.double().numpy()
        del ofvm, st
    err = rel_l2_cols(mine, res["fp32"])
    e64 = rel_l2_cols(mine, res["fp64"])
    o64 = rel_l2_cols(res["fp32"], res["fp64"])
    announce(f"  rel L2 vs fp32 oracle {err}; vs fp64 {e64} (fp32 oracle vs fp64 {o64})")
    # rho v starts at zero: after 10 steps the fp32 oracle is 1.5e-4 off fp64 on it
    # (measured: this build 7.9e-5 from the fp32 oracle, 1.26e-4 from fp64); the other
    # variables' fp32 oracle is within 1.1e-6 of fp64
    well = o64 <= 1e-5
    assert well[[0, 2, 3]].all(), o64
    assert np.all(err[well] <= 1e-4), err
    assert np.all(e64 <= np.maximum(o64, 1e-6)), (e64, o64)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("solver,extra,cfl", [("Butcher", "RK4", 0.5), ("AB3", "", 0.2)])
def test_2000_steps_1m_conservation(channel_1m, solver, extra, cfl):
    cfg, mesh, eq, U0, dt = _eq_1m(channel_1m, solver, extra, cfl)
    led = FluxLedger(eq.E_props.mesh, cfg)
    led.sample(eq.cells.state)
    ts = eq.t_solver
    for i in range(2000):
        ts._solve_step(None)
        led.sample(eq.cells.state, [dt])
        if i % 500 == 499:
            announce(f"  {solver}{extra} step {i + 1}")
    U = eq.cells.state
    assert torch.isfinite(U).all()
    d, integ, gross = led.report(U)
    err = np.abs(d - integ) / gross
    announce(f"{solver}{extra} 2000 steps, 1M cells, dt {dt:.3e}: total change {d[0]}, boundary-flux integral "
              f"{integ[0]}, |difference| / gross flux {err[0]}")
    assert np.all(err <= 1e-6), err


def _batch6():
    """Six meshes of the dataset generator, one above 100k cells."""
    from fvm_solver_amd.batch import random_channel_batch
    return random_channel_batch(6, seed=9, n_min=5000, n_max=200_000, device=DEV)


@pytest.mark.timeout(900)
def test_batch_against_oracle_per_mesh():
    from fvm_solver_amd.batch import BatchedSolver
    problems, params = _batch6()
    sizes = [p["cells"] for p in params]
    announce(f"batch of {sizes} cells: 20 RK4 steps, each mesh vs its fp32 oracle")
    assert max(sizes) >= 100_000
    bs = BatchedSolver(problems, device=DEV)
    dts = np.float32(bs.cfl_dt(0.5))
    bs.set_dt(dts)
    assert bs.E.fused
    for _ in range(20):
        bs.step()
    torch.set_num_threads(max(1, min(16, torch.get_num_threads())))
    for i, ((mesh, tags, cfg, U0), p) in enumerate(zip(problems, params)):
        pts, cells, fi, fb, tg = p["mesh_tuple"]
        _, o32 = O.from_mesh_tuple(pts, cells, fi, fb, tg, oracle_phys(cfg))
        st = O.OracleStepper(o32, "Butcher", "RK4", dt=float(dts[i]))
        U = U0.clone()
        for _ in range(20):
            U = st.step(U)
        err = rel_l2_cols(bs.state_of(i).cpu().numpy(), U.numpy())
        announce(f"  mesh {i} ({p['cells']} cells, Mach {p['mach']:.2f}, r {p['radius']:.3f}): rel L2 {err}")
        assert np.all(err <= 1e-4), (i, err)


@pytest.mark.timeout(900)
def test_batch64_long_run_conservation():
    from fvm_solver_amd.batch import BatchedSolver, random_channel_batch
    announce("generating the 64-mesh dataset batch")
    problems, params = random_channel_batch(64, seed=1, n_min=5000, n_max=200_000, device=DEV,
                                            progress=lambda i, n: announce(f"  mesh {i}: {n} cells") if i % 8 == 7
                                            else None)
    bs = BatchedSolver(problems, device=DEV)
    dts = np.float32(bs.cfl_dt(0.5))
    bs.set_dt(dts)
    cfg = problems[0][2]
    led = FluxLedger(bs.dm, cfg)
    announce(f"64-mesh batch, {bs.dm.N} cells: 200 RK4 steps")
    g = bs.capture_graph(2)
    led.sample(bs.state)
    for i in range(100):
        g.replay()
        led.sample(bs.state, 2 * dts.astype(np.float64))
    U = bs.state
    assert torch.isfinite(U).all()
    d, integ, gross = led.report(U)
    err = np.abs(d - integ) / gross
    worst = int(np.argmax(err.max(1)))
    announce(f"  worst mesh {worst}: |change - boundary integral| / gross {err[worst]}")
    assert np.all(err <= 3e-5), err.max(1)


@pytest.mark.timeout(900)
def test_channel3d_4m_against_oracle_and_conservation():
    """The 3-D channel at ~4M tetrahedra (the benchmark mesh): 500 RK4 steps at CFL 0.5 on
    the tile-fused stage, finite, with the conservation ledger (fp64 change of the totals
    against the trapezoid boundary-flux integral, to 1e-6 of the gross flux); and the
    state after 5 steps against 5 steps of the fp32 and the fp64 oracle.

    The transverse momenta start at zero and stay small, and the reference's fp32
    geometry (face normals from fp32 coordinate differences of cells ~1/200 of the
    domain) leaves residuals of ~1e-5 in its operators: the fp32 oracle itself is 9.3e-5
    off the fp64 oracle on rho v and rho w after 5 steps on this mesh (1.2e-4 to 2.3e-4
    after 2 to 20 steps on a 1M-tetrahedron mesh of the same family), while this build's
    operators come from fp64 geometry.  So the bar is: <= 1e-4 relative L2 against the
    fp32 oracle per conserved variable where that oracle is itself within 1e-5 of fp64
    (rho u, rho, E), and for every variable an error against the fp64 oracle no larger
    than the fp32 oracle's own (measured: 8.2e-5 and 8.6e-5 against its 9.3e-5 and
    9.4e-5 on rho v and rho w; <= 2.6e-8 on the others)."""
    from fvm_solver_amd.mesh_gen.channel import channel_mesh_3d
    announce("generating the 4M-tetrahedron channel mesh")
    pts, cells, (fi, fb), tags = channel_mesh_3d(4_000_000, seed=1)
    cfg = make_cfg(3, solver_name="Butcher", solver_extra="RK4")
    _, U0 = initial_state(pts[cells].mean(1), cfg, 3)
    announce(f"3-D channel: {len(cells)} tetrahedra, building the device layout")
    mesh, eq = product_equation(pts, cells, fi, fb, tags, cfg, us_init=U0)
    assert eq.E_props.fused, eq.E_props.mesh.tiles_reason
    dt = float(eq.cfl_dt(eq.cells.get_values()[0], 0.5))
    eq.t_solver.set_dt(dt)
    led = FluxLedger(eq.E_props.mesh, cfg)
    led.sample(eq.cells.state)
    ts = eq.t_solver
    K = 5
    UK = None
    for i in range(500):
        ts._solve_step(None)
        led.sample(eq.cells.state, [dt])
        if i == K - 1:
            UK = eq.cells.state.cpu().numpy()
        if i % 100 == 99:
            announce(f"  3-D step {i + 1}")
    U = eq.cells.state
    assert torch.isfinite(U).all()
    d, integ, gross = led.report(U)
    err = np.abs(d - integ) / gross
    announce(f"3-D 500 steps, dt {dt:.3e}: |total change - boundary-flux integral| / gross flux {err[0]}")
    assert np.all(err <= 1e-6), err
    del led, eq
    torch.set_num_threads(max(1, min(16, torch.get_num_threads())))
    res = {}
    for name, dtype in (("fp32", torch.float32), ("fp64", torch.float64)):
        announce(f"3-D: building the {name} oracle")
        _, o = O.from_mesh_tuple(pts, cells, fi, fb, tags, oracle_phys(cfg), dtype=dtype)
        st = O.OracleStepper(o, "Butcher", "RK4", dt=dt)
        Uo = U0.clone().to(dtype)
        for i in range(K):
            Uo = st.step(Uo)
            announce(f"  {name} oracle step {i + 1}")
        res[name] = Uo.double().numpy()
        del o, st, Uo
    e32 = rel_l2_cols(UK, res["fp32"])
    e64 = rel_l2_cols(UK, res["fp64"])
    o_e64 = rel_l2_cols(res["fp32"], res["fp64"])
    announce(f"3-D 4M, {K} RK4 steps: rel L2 vs fp32 oracle {e32}; vs fp64 oracle {e64}; "
             f"fp32 oracle vs fp64 oracle {o_e64}")
    well = o_e64 <= 1e-5
    assert well[[0, 3, 4]].all(), o_e64
    assert np.all(e32[well] <= 1e-4), e32
    assert np.all(e64 <= np.maximum(o_e64, 1e-6)), (e64, o_e64)
