"""The time loop's saves in the reference's npz format (ds_saving/saving.py), checked
against the CPU oracle stepped the same way: the saved cell primitives and the boundary
facets' two-side mean primitives, at every save time, after the reference's warm-up."""
import os

import numpy as np
import pytest
import torch

from fvm_solver_amd.ds_saving.saving import load_step
from oracle import fvm_oracle as O
from tests.gpu_helpers import make_cfg, oracle_phys, product_equation
from tests.golden_cases import load, mesh_tuple

pytestmark = pytest.mark.gpu


def test_saver_files_match_oracle(tmp_path):
    d = load("ref_2d.npz")
    pts, cells, fi, fb, tags = mesh_tuple(d)
    dt = float(d["step_dt"])
    save_dir = str(tmp_path / "saves")
    cfg = make_cfg(2, solver_name="Butcher", solver_extra="RK4", dt=dt, n_iter=6, save_t=2 * dt * 0.999,
                   save_dir=save_dir, print_i=3)
    U0 = torch.from_numpy(d["step_U0"])
    mesh, eq = product_equation(pts, cells, fi, fb, tags, cfg, us_init=U0)
    eq.t_solver.solve()
    files = sorted(f for f in os.listdir(save_dir) if f.startswith("t_"))
    mp = np.load(os.path.join(save_dir, "mesh_props.npz"), allow_pickle=False)
    assert set(mp.files) == {"triangles", "vertices", "centroids", "edges", "bc_midpoints", "bc_normals",
                             "bc_edge_masK", "bc_type_str"}
    assert mp["triangles"].shape == cells.shape and mp["bc_midpoints"].shape == (len(fb), 2)

    # oracle: warm-up step with the state restored (t_solvers.py:88-93), save at t = 0,
    # then the same steps, saving whenever t crosses a multiple of save_t
    _, o = O.from_mesh_tuple(pts, cells, fi, fb, tags, oracle_phys(cfg))
    st = O.OracleStepper(o, "Butcher", "RK4", dt=dt)
    U = U0.clone()
    st.step(U)
    expect = []
    bc_rows = torch.arange(len(fi), len(fi) + len(fb))       # boundary facets are stored last

    def snap(t, U):
        prim = o.state_to_primative(U)
        expect.append((t, prim.numpy(), o.last_face[bc_rows].mean(1).numpy()))

    snap(0.0, U)
    t, k = 0.0, 1
    for _ in range(6):
        U = st.step(U)
        t += dt
        if t >= k * cfg.save_t:
            k += 1
            snap(t, U)
    assert len(files) == len(expect), (files, [e[0] for e in expect])
    for (te, pe, be), f in zip(sorted(expect, key=lambda e: e[0]), sorted(files, key=lambda s: float(s[2:-4]))):
        ts_, cell, bc = load_step(os.path.join(save_dir, f))
        assert abs(float(ts_) - te) <= 1e-6 * max(te, dt)
        std = pe.std(0, ddof=1)
        ec = np.abs(cell.numpy() - pe).max(0) / np.maximum(std, 1e-8)
        eb = np.abs(bc.numpy() - be).max(0) / np.maximum(std, 1e-8)
        print(f"[gpu-test]   {f}: cell err/std {ec.max():.2e}, bc err/std {eb.max():.2e}")
        assert ec.max() < 2e-3 and eb.max() < 2e-3          # fp16 storage of normalised values


def test_exact_interval_lands_on_save_times(tmp_path):
    """exact_interval clamps the step that would cross a save time (t_solvers.py:112-119):
    saves land on k * save_t, and the clamped trajectory matches the oracle's."""
    d = load("ref_2d.npz")
    pts, cells, fi, fb, tags = mesh_tuple(d)
    dt = float(np.float32(d["step_dt"]))
    save_t = 2.5 * dt
    save_dir = str(tmp_path / "saves")
    cfg = make_cfg(2, solver_name="Butcher", solver_extra="RK4", dt=dt, n_iter=6, save_t=save_t,
                   save_dir=save_dir, exact_interval=True)
    U0 = torch.from_numpy(d["step_U0"])
    mesh, eq = product_equation(pts, cells, fi, fb, tags, cfg, us_init=U0)
    eq.t_solver.solve()
    times = sorted(float(load_step(os.path.join(save_dir, f))[0]) for f in os.listdir(save_dir) if f.startswith("t_"))
    print(f"[gpu-test]   save times / save_t: {[t / save_t for t in times]}")
    assert len(times) >= 3
    for k, t in enumerate(times):
        assert abs(t - k * save_t) <= 1e-5 * save_t
    # the same clamped sequence on the oracle
    _, o = O.from_mesh_tuple(pts, cells, fi, fb, tags, oracle_phys(cfg))
    st = O.OracleStepper(o, "Butcher", "RK4", dt=dt)
    U = U0.clone()
    st.step(U)
    t, nxt = 0.0, save_t
    for _ in range(6):
        h = dt
        if t + dt > nxt:
            h = np.float32(nxt - t)
            nxt += save_t
        st.dt.fill_(float(h))
        U = st.step(U)
        t += float(h)
        st.dt.fill_(dt)
    err = np.abs(eq.cells.state.cpu().double().numpy() - U.double().numpy()).max() / np.abs(U.numpy()).max()
    print(f"[gpu-test]   final state vs oracle (max rel) {err:.2e}")
    assert err < 1e-5


def test_profile_mode_writes_a_trace(tmp_path, monkeypatch):
    d = load("ref_2d.npz")
    pts, cells, fi, fb, tags = mesh_tuple(d)
    cfg = make_cfg(2, solver_name="Butcher", solver_extra="RK4", dt=float(d["step_dt"]), n_iter=4, profile=True)
    mesh, eq = product_equation(pts, cells, fi, fb, tags, cfg, us_init=torch.from_numpy(d["step_U0"]))
    monkeypatch.chdir(tmp_path)
    eq.t_solver.solve()
    assert os.path.getsize(tmp_path / "trace.json") > 0
