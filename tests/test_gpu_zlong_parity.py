"""Long-horizon parity on the 2-D channel at ~20k cells: 1000 fused RK4 steps on the GPU
against 1000 steps of the fp32 CPU oracle from the same state, with the error of both
against the fp64 oracle reported.

Tolerance: relative L2 <= 1e-4 per conserved variable after 1000
steps.  The fp64 run is reported, and bounds the product's error: the product may not
be further from fp64 than 10x the fp32 oracle's own distance to it (or 1e-5).
"""
import numpy as np
import pytest
import torch

from fvm_solver_amd.mesh_gen.channel import channel_mesh_2d
from oracle import fvm_oracle as O
from tests.gpu_helpers import DEV, initial_state, make_cfg, oracle_phys, product_equation, rel_l2_cols

pytestmark = pytest.mark.gpu

N_CELLS, STEPS = 20000, 1000


@pytest.mark.timeout(1500)
def test_rk4_1000_steps_20k_cells():
    torch.set_num_threads(max(1, min(16, torch.get_num_threads())))
    pts, cells, (fi, fb), tags = channel_mesh_2d(N_CELLS, seed=1)
    cfg = make_cfg(2, solver_name="Butcher", solver_extra="RK4")
    mesh, eq = product_equation(pts, cells, fi, fb, tags, cfg)
    prim0, U0 = initial_state(mesh.centroids.numpy(), cfg, 2)
    eq.cells.state = U0.to(DEV)
    dt = float(eq.cfl_dt(prim0.to(DEV), cfg.cfl))
    eq.t_solver.set_dt(dt)
    print(f"[gpu-test] 1000-step parity: N={mesh.n_cells} dt={dt:.4e}", flush=True)
    ts = eq.t_solver
    for _ in range(STEPS):
        ts._solve_step(None)
    mine = eq.cells.state.cpu().numpy()
    assert np.all(np.isfinite(mine))
    ph = oracle_phys(cfg)
    results = {}
    for name, dtype in (("fp32", torch.float32), ("fp64", torch.float64)):
        _, ofvm = O.from_mesh_tuple(pts, cells, fi, fb, tags, ph, dtype=dtype)
        st = O.OracleStepper(ofvm, "Butcher", "RK4", dt=dt)
        U = U0.to(dtype)
        for i in range(STEPS):
            U = st.step(U)
            if i % 250 == 249:
                print(f"[gpu-test]   oracle {name} step {i + 1}", flush=True)
        results[name] = U.double().numpy()
    e32 = rel_l2_cols(mine, results["fp32"])
    e64 = rel_l2_cols(mine, results["fp64"])
    o64 = rel_l2_cols(results["fp32"], results["fp64"])
    print(f"[gpu-test]   HIP vs fp32 oracle {e32}; HIP vs fp64 {e64}; fp32 oracle vs fp64 {o64}", flush=True)
    assert np.all(e32 <= 1e-4), e32
    assert np.all(e64 <= np.maximum(10 * o64, 1e-5)), (e64, o64)
