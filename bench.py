"""Benchmark: fused HIP RK4 step of the 2-D compressible NS channel at ~1M triangle cells.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cells N] [--no-cpu-baseline]

Workload: seeded Poisson-disk/Delaunay mesh of [0,2]x[0,1] minus a disk r=0.1 at
(0.5,0.5), ~1M cells, 4 fp32 conserved variables, freestream Mach 0.3 + Gaussian
density bump, Re 200, classical RK4 (Butcher "RK4") at CFL 0.5.  One step = 4 fused RHS
stages.  Prints one JSON line: metric = cell-updates/s (cells x steps / s), the dominant
kernel's HBM roofline (algorithmic bytes / measured duration), and the CPU oracle timed on
a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
import torch

HBM_PEAK_GBS = 8000.0


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", choices=("channel2d", "channel3d", "batch"), default="channel2d",
                    help="channel2d: the 1 M-cell 2-D channel (the headline); channel3d: the 3-D channel "
                         "(tetrahedra, C = 5; default 4 M cells); batch: 64 random 2-D channel meshes of "
                         "5k-200k cells in one block-diagonal layout, RK4 steps replayed as a HIP graph")
    ap.add_argument("--integrator", choices=("RK4", "AB3"), default="RK4",
                    help="fixed-dt classical RK4 (4 fused stages per step) or Adams-Bashforth 3 (one fused stage)")
    ap.add_argument("--cfl", type=float, default=None,
                    help="CFL number of the fixed time step (default 0.5 for RK4; 0.2 for AB3, whose stability "
                         "interval on the negative real axis is only [-6/11, 0])")
    ap.add_argument("--meshes", type=int, default=64, help="meshes of the batch configuration")
    ap.add_argument("--cells", type=int, default=None, help="cells (default 1 M in 2-D, 4 M in 3-D)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernel-timing-steps", type=int, default=20)
    ap.add_argument("--mesh-cache", default=None, help="npz file to reuse the generated mesh from (created if absent)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--fused-threads", type=int, default=None,
                    help="threads per tile of the tile-fused stage (256, 320, 384, 512) instead of the autotuner's "
                         "choice (with --no-autotune)")
    ap.add_argument("--traffic-file", default=None,
                    help="measured HBM bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated); default "
                         "profiles/rk4_channel2d_1M_traffic.json, rk4_channel3d_4M_traffic.json or "
                         "rk4_batch64_traffic.json by --config (used when its cell count matches the run)")
    ap.add_argument("--no-spmv", action="store_true",
                    help="skip timing the drop-in ELL SpMV/SpMM (ell_spmm, ell_spmv) on the reference's operators of "
                         "the mesh (G_mats interleaved, flux_mat, A_face_grad's pattern), reported as config.spmv")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one 256 MiB device-to-device copy (calibrates HBM byte counters)")
    g = ap.add_argument_group("learned corrector (--learned)")
    g.add_argument("--learned", action="store_true",
                   help="measure the learned window corrector against the plain RK4 solver instead")
    g.add_argument("--train-meshes", type=int, default=64)
    g.add_argument("--train-windows", type=int, default=157, help="windows per training mesh (64 x 157 = 10k pairs)")
    g.add_argument("--train-iters", type=int, default=12000,
                   help="Adam iterations (three layers: 2000 gave 420/1024 accepted predictions, 6000 639/1024; "
                        "one layer: 6000 610/1024, 12000 645/1024)")
    g.add_argument("--eval-meshes", type=int, default=64)
    g.add_argument("--eval-windows", type=int, default=16)
    g.add_argument("--eval-spinup", type=int, default=20,
                   help="windows of plain RK4 run on the evaluation meshes before the comparison starts")
    g.add_argument("--window", type=int, default=8, help="solver steps per predicted window")
    g.add_argument("--settle", type=int, default=1, help="plain RK4 steps after each window")
    g.add_argument("--layers", type=int, default=1,
                   help="message-passing layers of the corrector network (3: 1.03x plain, 2: 1.12x, 1: 1.26x at "
                        "6000 iterations; the input already carries one ring of least-squares gradients)")
    g.add_argument("--n-min", type=int, default=5000)
    g.add_argument("--n-max", type=int, default=200_000)
    g.add_argument("--oracle-meshes", type=int, default=2,
                   help="smallest evaluation meshes re-run by the FP64 CPU oracle (the mesh farthest from plain RK4 "
                        "is re-run as well)")
    g.add_argument("--tol", type=float, default=3.5e-4,
                   help="defect threshold of the check (3.5e-4 keeps every evaluation mesh's state within 1e-3 of "
                        "plain RK4 after 17 windows; 1e-3 lets the largest meshes drift to 2e-3)")
    ap.add_argument("--cpu-config1", action="store_true",
                    help="time only the CPU baseline of the 20k-cell 2-D channel, the full 1000 RK4 steps (the fp32 "
                         "torch-CPU restatement of the reference), on the usable physical cores and on one thread")
    g.add_argument("--corrector-timing", action="store_true",
                   help="time the pieces of one corrector window on the evaluation batch (untrained network: "
                        "the kernels' cost does not depend on the weights) against a plain window, then exit")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


_TRAFFIC = {"channel2d": "rk4_channel2d_1M_traffic.json", "channel3d": "rk4_channel3d_4M_traffic.json",
            "batch": "rk4_batch64_traffic.json"}


def traffic_of(a, kernel, cells):
    """Calibrated rocprofv3 HBM bytes per launch of ``kernel`` for this workload (read +
    write), or None when the file is absent or measured on a different cell count."""
    path = a.traffic_file or os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                                          _TRAFFIC[a.config])
    if not os.path.exists(path):
        return None, None
    tr = json.load(open(path))
    k = tr.get("kernels", {}).get(kernel)
    if not k or tr.get("cells") != cells:
        return None, None
    return k["read_bytes"] + k["write_bytes"], os.path.basename(path)


def host_cpu():
    """CPU model, physical cores (sockets x cores per socket, from lscpu) and the CPUs
    this process may run on."""
    import subprocess
    info = {"model": None, "physical_cores": None, "logical_cpus": os.cpu_count()}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        kv = dict(line.split(":", 1) for line in txt.splitlines() if ":" in line)
        kv = {k.strip(): v.strip() for k, v in kv.items()}
        info["model"] = kv.get("Model name")
        if kv.get("Core(s) per socket") and kv.get("Socket(s)"):
            info["physical_cores"] = int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["affinity_cpus"] = os.cpu_count()
    return info


def tile_summary(dm):
    """Tile tables of the fused stage: tiles, LDS per workgroup, halo cells (H1, H2)
    and foreign facets in total."""
    t = dm.tiles
    if t is None:
        return {"reason": dm.tiles_reason}
    return {"n": int(t["n_tiles"]), "lds_bytes": int(t["lds_bytes"]), "h1_cells": int(t["halo_cells"][0]),
            "h2_cells": int(t["halo_cells"][1]), "foreign_facets": int(t["foreign"]),
            "ncl": int(t["ncl"]), "ngl": int(t["ngl"]), "nfl": int(t["nfl"]),
            "threads_per_tile": dm.fused_threads}


def facet_kernel_bytes(N, F, Fb, D):
    """Algorithmic (compulsory) bytes of one k_facet_flux launch, every array read or
    written once (fp32 values, int32 indices).  Per facet: its two cells (2 indices),
    the geometry record (n, n_hat, |n|, 1/(n.d), X_orthog, both centroid->midpoint
    displacements) and the flux record written; per cell: prim and the gradient record
    ((D+1)C floats); per boundary facet: the boundary values and the two face-gradient
    coefficients."""
    C = D + 2
    geo = 12 if D == 2 else 20
    return F * 4 * (2 + geo + C) + N * 4 * (C + (D + 1) * C) + Fb * 4 * (3 * C)


def div_kernel_bytes(N, F, D, stage):
    """k_div (stage epilogue): per cell the facet list, info word, 1/vol, U0, the RK
    accumulator (read after stage 0), the accumulator or new state written, and the
    next-stage primitives written; per facet its flux record, read once."""
    C, NL = D + 2, D + 1
    return N * 4 * (NL + 1 + 1 + C + (C if stage > 0 else 0) + 2 * C) + F * 4 * C


def model_stage_bytes(N, F, Fb, D, ab3=False):
    """Algorithmic bytes of one RHS stage of a single-pass fused kernel, by the per-unit
    model the roofline is quoted on (every listed array read or written once, fp32
    values, int32 indices): per cell 99 B (2-D) / 164 B (3-D) of static data (neighbours,
    LSQ row with its diagonal, facet list, flux_mat values, slots, displacements, eps);
    per interior facet 28 / 36 B (cells, normal, 1/(n.d), X_orthog); per boundary facet
    46 / 50 B; the state traffic of classical RK4 averaged over its four stages, 17 C
    bytes per cell, 68 / 85 B (Adams-Bashforth 3: 20 C bytes, 80 / 100 B)."""
    C = D + 2
    static, fi, fb = (99, 28, 46) if D == 2 else (164, 36, 50)
    dyn = 20 * C if ab3 else 17 * C
    return N * (static + dyn) + (F - Fb) * fi + Fb * fb


def fused_kernel_bytes(N, F, Fb, D, stage, tiles, ab3=False):
    """k_fused (tile-fused stage), every array this implementation reads or writes once:
    per cell prim, the LSQ values of the neighbour columns (the own column is derived),
    the displacements of all slots but the last (which is minus their sum), eps, one
    packed word per slot (local neighbour, list position, side, LSQ order), 1/vol and the
    stage epilogue's traffic (as k_div); per facet the compact geometry (n, 1/(n.d),
    X_orthog's D - 1 tangent coordinates) and the descriptor; per boundary facet its BC
    data, boundary value written and the interior cell's gradient record read and
    written; the tile table.  Halo re-reads (tile + H1 + H2 staging) are not counted:
    they show in the measured traffic."""
    C, NL = D + 2, D + 1
    per_cell = C + NL * D + (NL - 1) * D + 1 + NL + 1
    # RK stage: U0, the accumulator after stage 0, accumulator (or state) and next-stage
    # primitives written; Adams-Bashforth 3: U0 and two past RHS read, state, RHS and
    # primitives written
    dyn = 6 * C if ab3 else C + (C if stage > 0 else 0) + 2 * C
    per_facet = 2 * D + 1
    per_bc = 1 + C + D + 2 * C + (2 * D + 3) + 2 + C + 2 * (D + 1) * C
    return 4 * (N * (per_cell + dyn) + F * per_facet + Fb * per_bc + tiles * 12)


def spmv_operators(mesh, C):
    """The reference's sparse operators of a mesh, as torch CSR on the host: G_mats with
    rows interleaved (row d cell + dir; sparse.py:281-325, fvm_mesh.py:5-42), flux_mat
    (value (2 slot - 1) / vol = -sign / vol on each cell's facets; facet_process.py:200-226)
    and A_face_grad's pattern (rows f C + c; interior facets -1/(n.d), +1/(n.d) on the two
    cells' component c, boundary facets one entry; facet_process.py:114-198), with the
    right-hand sides they multiply: [N + F_b, C], [F, C], [C N]."""
    g = mesh.geo64
    N, F, Fb, D = mesh.n_cells, mesh.n_facets, mesh.n_bc_facet, mesh.dim
    _, _, G_mats, _ = mesh.cell_grad_stuff
    r, c, v = [], [], []
    for i, G in enumerate(G_mats):
        G = G.coalesce()
        r.append(G.indices()[0] * D + i)
        c.append(G.indices()[1])
        v.append(G.values())
    Gi = torch.sparse_coo_tensor(torch.stack([torch.cat(r), torch.cat(c)]), torch.cat(v), (D * N, N + Fb))
    c2f, signs = torch.from_numpy(np.asarray(g["c2f"], np.int64)), torch.from_numpy(np.asarray(g["signs"], np.int64))
    vol = torch.from_numpy(np.asarray(g["volumes"])).float()
    rows = torch.arange(N)[:, None].expand_as(c2f)
    fm = torch.sparse_coo_tensor(torch.stack([rows.reshape(-1), c2f.reshape(-1)]),
                                 (-signs.float() / vol[:, None]).reshape(-1), (N, F))
    left, right = np.asarray(g["left"], np.int64), np.asarray(g["right"], np.int64)
    bcm = mesh.bc_facet_mask.numpy()
    fi = np.flatnonzero(~bcm)
    fb = np.flatnonzero(bcm)
    cen, nh = g["centroids"], g["normals_hat"]
    inv_d = 1.0 / np.einsum("nd,nd->n", nh[fi], cen[right[fi]] - cen[left[fi]])
    comp = np.arange(C)
    ri = (fi[:, None] * C + comp).reshape(-1)
    rl = (left[fi][:, None] * C + comp).reshape(-1)
    rr = (right[fi][:, None] * C + comp).reshape(-1)
    vi = np.repeat(inv_d, C)
    rb = (fb[:, None] * C + comp).reshape(-1)
    cb = (left[fb][:, None] * C + comp).reshape(-1)
    A = torch.sparse_coo_tensor(torch.from_numpy(np.stack([np.concatenate([ri, ri, rb]), np.concatenate([rl, rr, cb])])),
                                torch.from_numpy(np.concatenate([-vi, vi, np.ones(len(rb))])).float(), (C * F, C * N))
    return {"G_mats": (Gi.coalesce().to_sparse_csr(), (N + Fb, C)), "flux_mat": (fm.coalesce().to_sparse_csr(), (F, C)),
            "A_face_grad": (A.coalesce().to_sparse_csr(), (C * N,))}


def spmv_bench(mesh, C, dev, reps=50, cold_bytes=768e6, max_copies=16):
    """Per operator: ELL packing (csr_to_ell), the launch-shape autotuning of its first
    product, then products timed with stream events; GB/s from the algorithmic bytes
    nnz (4 B value + 4 B column) + the right-hand side's used columns + the output (the
    SpMV figure of the reference's operators).

    ``us`` (cold): the products cycle through R copies of the operator, right-hand side
    and output, R x working set >= 768 MB (three times the 256 MB MALL, at most 16
    copies), so each product's operands were last touched R - 1 products earlier and
    have been evicted from the MALL: HBM rates.  ``us_hot``: reps products on one copy
    back to back (a working set under 256 MB can stay in the MALL between them).
    Both launch the tuned shape through the C ABI directly."""
    from fvm_solver_amd import _native as nat
    from fvm_solver_amd.utils import ell_kernel as ek
    out = {}
    g = torch.Generator().manual_seed(7)
    stream = nat.stream_ptr(torch.device(dev))
    for name, (A, xshape) in spmv_operators(mesh, C).items():
        vals, cols = ek.csr_to_ell(A.to(dev))
        x = torch.randn(*xshape, generator=g).to(dev)
        fn = ek.ell_spmv if len(xshape) == 1 else ek.ell_spmm
        y = fn(vals, cols, x)                          # packs, autotunes, checks
        ref = (A.to_dense() @ x.cpu()) if A.shape[0] * A.shape[1] < 2 ** 26 else None
        B = 1 if len(xshape) == 1 else xshape[1]
        M, K = vals.shape
        nnz = int(A.values().numel())
        ncols = int(torch.unique(A.col_indices()).numel())
        nbytes = nnz * 8 + ncols * B * 4 + M * B * 4
        tuned = ek.best_config(M, B, int(K), False)
        shape = tuned["shape"] if tuned is not None else (1, 256)
        R = int(min(max_copies, max(2, -(-cold_bytes // nbytes))))
        copies = []
        for r in range(R):
            v = vals if r == 0 else vals.t().contiguous().t()      # the same slot-major layout
            c = cols if r == 0 else cols.t().contiguous().t()
            x2 = (x if r == 0 else x.clone()).reshape(x.shape[0], -1)
            yr = torch.empty(M, B, device=dev)
            copies.append(((nat.ptr(v), v.stride(0), v.stride(1), nat.ptr(c), 0, c.stride(0), c.stride(1), M, K,
                            nat.ptr(x2), x2.stride(0), x2.stride(1), B, 0, nat.ptr(yr), yr.stride(0), yr.stride(1)),
                           (v, c, x2, yr)))
        torch.cuda.synchronize()

        def timed(seq):
            for a, _ in seq[:2]:
                ek._launch(a, shape, stream)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for a, _ in seq:
                ek._launch(a, shape, stream)
            e1.record()
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) * 1e3 / len(seq)

        n_cold = R * max(1, -(-reps // R))
        us_cold = timed([copies[i % R] for i in range(n_cold)])
        us_hot = timed([copies[0]] * reps)
        ok = all(torch.equal(cp[1][3], y.reshape(M, B)) for cp in copies)
        out[name] = {"rows": M, "K": int(K), "nnz": nnz, "B": B, "us": us_cold, "us_hot": us_hot,
                     "algorithmic_bytes": nbytes, "gbs": nbytes / (us_cold * 1e-6) / 1e9,
                     "gbs_hot": nbytes / (us_hot * 1e-6) / 1e9, "copies": R, "copies_equal": ok,
                     "launch_shape": None if tuned is None else
                     {"rows_per_lane": shape[0], "threads_per_block": shape[1], "candidates_us": tuned["us"]},
                     "working_set_mb": nbytes / 1e6}
        if ref is not None:
            out[name]["rel_err"] = float((y.cpu() - ref).norm() / ref.norm())
        del copies
    return out


def block_kernel_us(E, stages, parts, reps, lead=2):
    """Average launch duration (us) of each stage kernel: one event pair around `reps`
    passes over `stages` (the FvmStage structs of one step) launching that kernel alone,
    back to back on the solver's stream, behind `lead` untimed passes that keep the queue
    full when the first event is processed (an event pair around a single launch, or
    a start event on an idle GPU, adds several us of command processing and launch
    latency per launch).  The kernels read the step's own buffers; their results are
    discarded (the state is not advanced).  The boundary-mean pre-pass is not timed: a
    stage's epilogue provides the next stage's partial sums, so the step loop never
    launches it."""
    out = {}
    for mask, name in parts.items():
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for _ in range(lead):
            for st in stages:
                E.stage(st, parts=mask)
        e0.record()
        for _ in range(reps):
            for st in stages:
                E.stage(st, parts=mask)
        e1.record()
        torch.cuda.synchronize()
        out[name] = 1e3 * e0.elapsed_time(e1) / (reps * len(stages))
    return out


def grad_kernel_bytes(N, Fb, D):
    """k_grad: prim, neighbour list, info, LSQ values ((NL+1) D), displacements, eps;
    writes the gradient record; boundary values read once."""
    C, NL = D + 2, D + 1
    return N * 4 * (C + NL + 1 + (NL + 1) * D + NL * D + 1 + (D + 1) * C) + Fb * 4 * C


def main():
    a = _parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}): batches shard by meshes and a large mesh "
              "by slabs of cells with a halo exchange per stage, but no multi-GPU figure is measured",
              file=sys.stderr)
        sys.exit(2)
    from fvm_solver_amd import _native as nat
    from fvm_solver_amd.config_fvm import ConfigChannel, ConfigChannel3D
    from fvm_solver_amd.fvm_equation import FluidConstitution2D, FluidConstitution3D, FVMEquation
    from fvm_solver_amd.mesh_gen.channel import channel_mesh_2d, channel_mesh_3d
    from fvm_solver_amd.mesh_utils.fvm_mesh import mesh_from_tuple
    from fvm_solver_amd.mesh_utils.mesh_store import channel_bc_tags

    D = 3 if a.config == "channel3d" else 2
    C = D + 2
    if a.cells is None:
        a.cells = 4_000_000 if D == 3 else 1_000_000

    if not torch.cuda.is_available():
        raise RuntimeError("bench.py needs a GPU")
    dev = torch.device("cuda:0")
    t0 = time.time()
    if a.mesh_cache and os.path.exists(a.mesh_cache):
        z = np.load(a.mesh_cache, allow_pickle=False)
        pts, cells, fi, fb, tags = z["pts"], z["cells"], z["fi"], z["fb"], [str(t) for t in z["tags"]]
    else:
        gen = channel_mesh_3d if D == 3 else channel_mesh_2d
        pts, cells, (fi, fb), tags = gen(a.cells, seed=a.seed)
        if a.mesh_cache:
            np.savez(a.mesh_cache, pts=pts, cells=cells, fi=fi, fb=fb, tags=np.array(tags))
    log(f"mesh: {len(cells)} cells, {len(fi) + len(fb)} facets ({len(fb)} boundary) in {time.time() - t0:.1f}s")
    Cfg = ConfigChannel3D if D == 3 else ConfigChannel
    ab3 = a.integrator == "AB3"
    cfg = Cfg(device="cuda:0", n_cells=a.cells, solver_name="AB3" if ab3 else "Butcher",
              solver_extra="" if ab3 else "RK4")
    t0 = time.time()
    mesh = mesh_from_tuple(pts, cells, fi, fb, device=dev)
    phy = FluidConstitution3D(cfg) if D == 3 else FluidConstitution2D(cfg)
    N, F, Fb = mesh.n_cells, mesh.n_facets, mesh.n_bc_facet
    c = mesh.geo64["centroids"]
    prim = np.zeros((N, C))
    prim[:, 0] = cfg.inlet_cfg.v_n_inf
    r2 = ((c - np.asarray(cfg.bump_centre)[:D]) ** 2).sum(1)
    prim[:, D] = 1.0 + cfg.bump_amp * np.exp(-r2 / (2 * cfg.bump_sigma ** 2))
    prim[:, D + 1] = cfg.inlet_cfg.T_inf
    prim = torch.from_numpy(prim).float()
    V, rho, T = prim[:, :D], prim[:, D:D + 1], prim[:, D + 1:]
    U0 = torch.cat(phy.primatives_to_state(V, rho, T), -1)
    eq = FVMEquation(cfg, phy, mesh, C, channel_bc_tags(tags, D), us_init=U0)
    cfl = a.cfl if a.cfl is not None else (0.2 if ab3 else cfg.cfl)
    dt = float(eq.cfl_dt(prim.to(dev), cfl))
    eq.t_solver.set_dt(dt)
    log(f"set-up {time.time() - t0:.1f}s, dt={dt:.4e}")
    ts = eq.t_solver

    for _ in range(a.warmup):
        ts._solve_step(None)
    tuned = None
    if not a.no_autotune:
        from fvm_solver_amd import autotune
        tuned = autotune.tune(lambda: ts._solve_step(None), eq.E_props, N, D, reps=10 if D == 2 else 4)
        log(f"autotune: {tuned['choice']} " + ", ".join(f"{k}:{v:.0f}us" for k, v in tuned["us_per_step"].items()))
    elif eq.E_props.mesh.tiles is None:
        log(f"tile-fused stage unavailable: {eq.E_props.mesh.tiles_reason}")
    elif a.fused_threads:
        eq.E_props.mesh.set_fused_threads(a.fused_threads)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    ev0.record()
    for _ in range(a.steps):
        ts._solve_step(None)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    gpu_ms = ev0.elapsed_time(ev1)
    ms_per_step = gpu_ms / a.steps
    value = N * a.steps / (gpu_ms / 1e3)
    state = eq.cells.state
    finite = bool(torch.isfinite(state).all())
    log(f"{a.steps} steps: {gpu_ms:.2f} ms GPU ({wall * 1e3:.2f} ms wall), {ms_per_step * 1e3:.1f} us/step, finite={finite}")

    if a.calib_copy:
        src = torch.empty(64 * 2 ** 20, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 256 MiB read + 256 MiB written")

    # ---- kernel durations ----
    # The step of the tile-fused pipeline is exactly `stages` launches of k_fused (the
    # boundary means of each stage come from the previous stage's epilogue), so the
    # average launch duration is the timed region divided by its launches: the figure the
    # roofline uses, and one a kernel trace of the same command must agree with (the
    # trace excludes the few-us dispatch gaps between launches, so it reads slightly less).
    E = eq.E_props
    parts = {16: "k_fused"} if E.fused else {2: "k_grad", 4: "k_facet_flux", 8: "k_div"}
    n_tiles = (N + 255) // 256
    s = ts.stages
    div_bytes = [div_kernel_bytes(N, F, D, i) for i in range(s)]
    fused_bytes = [fused_kernel_bytes(N, F, Fb, D, i, n_tiles, ab3=ab3) for i in range(s)]
    launches = a.steps * s
    step_us = 1e3 * gpu_ms / launches          # average duration of one stage launch (fused)
    if E.fused:
        avg_us = {"k_fused": step_us}
    else:
        U = eq.cells.state
        cur = ts._stage_prim0(U)
        out = ts._out[1] if U.data_ptr() == ts._out[0].data_ptr() else ts._out[0]
        flip = ts._flip
        stages = []
        for i in range(s):
            nxt = ts._prim[1 - flip]
            st = nat.FvmStage(prim=cur.data_ptr(), U0=U.data_ptr(), acc=ts._acc.data_ptr(), acc2=None, kbuf=None,
                              prim_next=nxt.data_ptr(), U_out=out.data_ptr(), dt=ts.dt.data_ptr(), stage=i,
                              last=int(i == s - 1), store_k=0, b=float(ts.b[i]), b2=float(ts.b2[i]))
            if i < s - 1:
                terms = [(j, float(ts.A[i + 1][j])) for j in range(i + 1) if ts.A[i + 1][j] != 0]
                st.n_inc = len(terms)
                for q, (j, cc) in enumerate(terms):
                    st.inc_idx[q], st.inc_coef[q] = j, cc
            stages.append(st)
            flip = 1 - flip
            cur = nxt
        # three kernels per stage: per-kernel durations from back-to-back launches of each
        # (informational); the roofline below is the stage's, from the timed region
        avg_us = block_kernel_us(E, stages, parts, a.kernel_timing_steps)
    log("kernel averages (us): " + ", ".join(f"{k}={v:.1f}" for k, v in avg_us.items()))
    kbytes = {"k_grad": float(grad_kernel_bytes(N, Fb, D)), "k_facet_flux": float(facet_kernel_bytes(N, F, Fb, D)),
              "k_div": float(np.mean(div_bytes)) if div_bytes else 0.0, "k_fused": float(np.mean(fused_bytes))}
    kgbs = {k: kbytes[k] / (avg_us[k] * 1e-6) / 1e9 for k in avg_us}
    # single-pass roofline of the step: a stage's algorithmic bytes by the per-unit model
    # (model_stage_bytes) over the measured time of one stage in the timed region; the
    # three-kernel pipeline moves more (the gradient records and facet fluxes make a
    # round trip through HBM), which this figure charges to it
    stage_bytes = float(model_stage_bytes(N, F, Fb, D, ab3=ab3))
    dom = "k_fused" if E.fused else max(avg_us, key=avg_us.get)
    dom_bytes = stage_bytes if E.fused else kbytes[dom]
    achieved = stage_bytes / (step_us * 1e-6) / 1e9
    log(f"stage: {stage_bytes / 1e6:.1f} MB algorithmic ({stage_bytes / N:.1f} B/cell, model) in {step_us:.1f} us "
        f"-> {achieved:.0f} GB/s ({achieved / HBM_PEAK_GBS:.3f} of peak); this implementation's own bytes "
        f"{kbytes['k_fused'] / 1e6:.1f} MB; per kernel GB/s: " + ", ".join(f"{k}={v:.0f}" for k, v in kgbs.items()))

    out = {
        "metric": f"cell-updates/s (fused HIP {a.integrator} step, {D}-D NS channel)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Poisson-disk Delaunay mesh, freestream + Gaussian bump)",
        "config": {"workload": ("3-D compressible NS channel with spherical obstacle (tetrahedra)" if D == 3 else
                                "2-D compressible NS channel with circular obstacle") + f", {a.integrator} at CFL {cfl:g}",
                   "cells": N, "facets": F, "boundary_facets": Fb,
                   "integrator": "AB3 (fixed dt)" if ab3 else "Butcher RK4",
                   "dt": dt, "finite": finite, "pipeline": "tile-fused stage" if E.fused else "three kernels",
                   "kernel_avg_us": avg_us,
                   "kernel_timing": ("timed region / launches (a step is `stages` k_fused launches)" if E.fused else
                                     "one event pair around back-to-back launches of each kernel"),
                   "kernel_implementation_bytes": {k: kbytes[k] for k in avg_us}, "kernel_gbs": kgbs,
                   "stage_us": step_us, "stage_algorithmic_bytes": stage_bytes,
                   "algorithmic_bytes_per_cell": stage_bytes / N,
                   "algorithmic_model": "per cell static 99 B (2-D) / 164 B (3-D) + per interior facet 28 / 36 B + "
                                        "per boundary facet 46 / 50 B + RK4 state traffic 17 C B per cell "
                                        "(AB3: 20 C B)"},
        "roofline": {"bound": "hbm", "kernel": dom if E.fused else "stage (k_grad + k_facet_flux + k_div)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "timing": "timed region: ms_per_step / stages per step",
                     "bytes_per_launch": stage_bytes},
    }
    if D == 3:
        hc = eq.E_props.mesh.h64
        out["config"]["cell_size"] = {"min_h": float(hc.min()), "mean_h": float(hc.mean()),
                                      "min_over_mean": float(hc.min() / hc.mean())}
    out["config"]["persist_blocks"] = tuned["persist_blocks"] if tuned else 0
    out["config"]["tiles"] = tile_summary(E.mesh)
    if tuned:
        out["config"]["autotune_us_per_step"] = tuned["us_per_step"]
    out["config"]["algorithmic_bytes_per_launch"] = dom_bytes
    if not ab3 and E.fused:
        t, src = traffic_of(a, dom, N)
        if t is not None:
            out["roofline"]["traffic"] = t
            out["roofline"]["traffic_source"] = src

    if not a.no_spmv and D == 2:          # (the 3-D operators take minutes to assemble on the host)
        sp = spmv_bench(mesh, C, dev)
        out["config"]["spmv"] = sp
        log("ELL products: " + ", ".join(f"{k} {v['us']:.1f} us {v['gbs']:.0f} GB/s" for k, v in sp.items()))

    if not a.no_cpu_baseline:
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        from oracle import fvm_oracle as O
        ins, ex = cfg.inlet_cfg, cfg.exit_cfg
        ph = O.Phys(T_0=cfg.T_0, viscosity=cfg.viscosity, visc_bulk=cfg.visc_bulk, thermal_cond=cfg.thermal_cond,
                    S_const=cfg.S_const, gamma=cfg.gamma, C_v=cfg.C_v, v_factor=cfg.v_factor, lim_p=cfg.lim_p,
                    lim_K=cfg.lim_K, inlet=O.BCSpec(ins.v_n_inf, ins.v_t_inf, ins.rho_inf, ins.T_inf, ins.v_inf),
                    exit=O.BCSpec(ex.v_n_inf, ex.v_t_inf, ex.rho_inf, ex.T_inf, ex.v_inf))
        cpu = host_cpu()
        # the physical cores this process may use (the box grants a share of the host)
        threads = min(x for x in (cpu["physical_cores"], cpu["affinity_cpus"], cpu["logical_cpus"]) if x)
        _, ofvm = O.from_mesh_tuple(pts, cells, fi, fb, tags, ph)
        if D == 3:
            a.cpu_steps = max(1, a.cpu_steps // 6)
        if ab3:
            a.cpu_steps *= 4             # one RHS per step

        def cpu_rate(n_threads, n_steps):
            torch.set_num_threads(n_threads)
            stepper = O.OracleStepper(ofvm, "AB3" if ab3 else "Butcher", "" if ab3 else "RK4", dt=dt)
            Uc = stepper.step(U0.clone())            # warm-up
            t0 = time.perf_counter()
            for _ in range(n_steps):
                Uc = stepper.step(Uc)
            return N * n_steps / (time.perf_counter() - t0), time.perf_counter() - t0

        rate, cpu_s = cpu_rate(threads, a.cpu_steps)
        n1 = max(1, a.cpu_steps // (threads * 2))
        rate1, cpu1_s = cpu_rate(1, n1)
        out["cpu_baseline"] = {"value": rate, "unit": "cell-updates/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{a.cpu_steps} {a.integrator} steps of the same {N}-cell mesh (after 1 warm-up "
                                         f"step), fp32 torch-CPU restatement of the reference, {cpu_s:.1f}s on {threads} "
                                         f"threads; single thread: {n1} steps, {cpu1_s:.1f}s",
                               "single_thread_value": rate1, "cpu_model": cpu["model"],
                               "host_physical_cores": cpu["physical_cores"], "affinity_cpus": cpu["affinity_cpus"]}
        log(f"cpu baseline ({cpu['model']}, {cpu['physical_cores']} physical cores, {cpu['affinity_cpus']} usable): "
            f"{cpu_s:.1f}s for {a.cpu_steps} steps on {threads} threads; 1 thread {cpu1_s:.1f}s for {n1} steps")
    print(json.dumps(out), flush=True)


def batch_main(a):
    """Dataset-generation throughput: a batch of independent channel meshes stepped
    together (one launch per kernel for the whole batch), RK4 steps captured as a HIP
    graph; per-kernel durations from part-wise launches of the same stages."""
    from fvm_solver_amd import _native as nat
    from fvm_solver_amd.batch import BatchedSolver, random_channel_batch
    dev = torch.device("cuda:0")
    t0 = time.time()
    problems, params = random_channel_batch(a.meshes, seed=a.seed, n_min=5000, n_max=200_000, device=dev)
    s = BatchedSolver(problems, device=dev)
    dts = s.cfl_dt(0.5)
    s.set_dt(dts)
    N, F, Fb = s.dm.N, s.dm.F, s.dm.Fb
    log(f"batch: {a.meshes} meshes, {N} cells, {F} facets ({Fb} boundary), set-up {time.time() - t0:.1f}s")
    for _ in range(a.warmup):
        s.step()
    if not a.no_autotune:
        from fvm_solver_amd import autotune
        tuned = autotune.tune(s.step, s.E, N, 2, reps=6)
        log(f"autotune: {tuned['choice']} " + ", ".join(f"{k}:{v:.0f}us" for k, v in tuned["us_per_step"].items()))
    elif a.fused_threads and s.E.fused:
        s.E.mesh.set_fused_threads(a.fused_threads)
    n_graph = 10
    g = s.capture_graph(n_graph)
    steps = max(n_graph, (a.steps // n_graph) * n_graph)
    g.replay()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(steps // n_graph):
        g.replay()
    ev1.record()
    torch.cuda.synchronize()
    gpu_ms = ev0.elapsed_time(ev1)
    ms_per_step = gpu_ms / steps
    value = N * steps / (gpu_ms / 1e3)
    finite = bool(torch.isfinite(s.state).all())
    log(f"{steps} steps (graph of {n_graph}): {gpu_ms:.2f} ms, {ms_per_step * 1e3:.1f} us/step, finite={finite}")
    if a.calib_copy:
        src = torch.empty(64 * 2 ** 20, dtype=torch.float32, device=dev).fill_(1.0)
        torch.empty_like(src).copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 256 MiB read + 256 MiB written")
    # kernel durations: a replayed step is `stages` k_fused launches, so the average launch
    # is the timed region over its launches; the three-kernel pipeline's kernels are
    # timed by back-to-back launches of each
    E = s.E
    parts = {16: "k_fused"} if E.fused else {2: "k_grad", 4: "k_facet_flux", 8: "k_div"}
    step_us = 1e3 * gpu_ms / (steps * s.stages)
    s._ensure_prim()
    U0, out, cur, flip = s._U[s._cur], s._U[1 - s._cur], s._prim[s._flip], s._flip
    stages = []
    for i in range(s.stages):
        nxt = s._prim[1 - flip]
        st = nat.FvmStage(prim=cur.data_ptr(), U0=U0.data_ptr(), acc=s._acc.data_ptr(), acc2=None, kbuf=None,
                          prim_next=nxt.data_ptr(), U_out=out.data_ptr(), dt=s.dt.data_ptr(), stage=i,
                          last=int(i == s.stages - 1), store_k=0, b=float(s.b[i]), b2=float(s.b2[i]))
        if i < s.stages - 1 and s.A[i + 1][i] != 0:
            st.n_inc, st.inc_idx[0], st.inc_coef[0] = 1, i, float(s.A[i + 1][i])
        stages.append(st)
        flip = 1 - flip
        cur = nxt
    avg_us = {"k_fused": step_us} if E.fused else block_kernel_us(E, stages, parts, 10)
    kbytes = {"k_grad": float(grad_kernel_bytes(N, Fb, 2)), "k_facet_flux": float(facet_kernel_bytes(N, F, Fb, 2)),
              "k_div": float(np.mean([div_kernel_bytes(N, F, 2, i) for i in range(4)])),
              "k_fused": float(np.mean([fused_kernel_bytes(N, F, Fb, 2, i, (N + 255) // 256) for i in range(4)]))}
    kgbs = {k: kbytes[k] / (avg_us[k] * 1e-6) / 1e9 for k in avg_us}
    dom = max(avg_us, key=avg_us.get)
    stage_bytes = float(model_stage_bytes(N, F, Fb, 2))
    achieved = stage_bytes / (step_us * 1e-6) / 1e9
    log("kernel averages (us): " + ", ".join(f"{k}={v:.1f}" for k, v in avg_us.items()))
    log(f"stage: {stage_bytes / 1e6:.1f} MB algorithmic (model) in {step_us:.1f} us -> {achieved:.0f} GB/s "
        f"({achieved / HBM_PEAK_GBS:.3f} of peak)")
    out = {
        "metric": "cell-updates/s (batched 2-D NS channels, HIP-graph RK4)",
        "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": steps, "warmup": a.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded random channel meshes, freestream + Gaussian bump)",
        "config": {"workload": f"{a.meshes} independent 2-D channel meshes (cells log-uniform 5k-200k, obstacle "
                               "x ~ U(0.3, 0.8), r ~ U(0.05, 0.15), Mach ~ U(0.1, 0.6)), RK4 at CFL 0.5",
                   "cells": N, "facets": F, "boundary_facets": Fb, "finite": finite, "graph_steps": n_graph,
                   "pipeline": "tile-fused stage" if E.fused else "three kernels",
                   "tiles": tile_summary(E.mesh), "kernel_avg_us": avg_us,
                   "kernel_timing": ("timed region / launches (a step is 4 k_fused launches)" if E.fused else
                                     "one event pair around back-to-back launches of each kernel"),
                   "kernel_gbs": kgbs, "stage_us": step_us, "kernel_implementation_bytes": kbytes,
                   "algorithmic_bytes_per_cell": stage_bytes / N},
        "roofline": {"bound": "hbm", "kernel": "k_fused" if E.fused else "stage (three kernels)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "timing": "timed region: ms_per_step / stages per step",
                     "bytes_per_launch": stage_bytes},
    }
    if E.fused:
        t, src = traffic_of(a, "k_fused", N)
        if t is not None:
            out["roofline"]["traffic"] = t
            out["roofline"]["traffic_source"] = src
    print(json.dumps(out), flush=True)


def _oracle_phys(cfg):
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import fvm_oracle as O
    ins, ex = cfg.inlet_cfg, cfg.exit_cfg
    return O.Phys(T_0=cfg.T_0, viscosity=cfg.viscosity, visc_bulk=cfg.visc_bulk, thermal_cond=cfg.thermal_cond,
                  S_const=cfg.S_const, gamma=cfg.gamma, C_v=cfg.C_v, v_factor=cfg.v_factor, lim_p=cfg.lim_p,
                  lim_K=cfg.lim_K, inlet=O.BCSpec(ins.v_n_inf, ins.v_t_inf, ins.rho_inf, ins.T_inf, tuple(ins.v_inf)),
                  exit=O.BCSpec(ex.v_n_inf, ex.v_t_inf, ex.rho_inf, ex.T_inf, tuple(ex.v_inf)))


def learned_main(a):
    """Learned corrector vs plain RK4 on a batch of 2-D channel meshes: simulated time
    per wall second, acceptance, and error against the FP64 CPU oracle (the checker)."""
    from fvm_solver_amd.batch import BatchedSolver, random_channel_batch
    from fvm_solver_amd.learned.corrector import LearnedCorrector
    dev = torch.device("cuda:0")
    t0 = time.time()
    tp, tparams = random_channel_batch(a.train_meshes, seed=a.seed + 1000, n_min=a.n_min, n_max=a.n_max, device=dev)
    ts = BatchedSolver(tp, device=dev)
    ts.set_dt(ts.cfl_dt(0.5))
    for _ in range(a.eval_spinup * a.window):         # past the impulsive start, as the evaluation
        ts.step()
    lc = LearnedCorrector(ts, k=a.window, layers=a.layers, seed=a.seed, tol=a.tol, settle=a.settle)
    data = lc.generate(a.train_windows, log=log)
    gen_s = time.time() - t0
    log(f"training data: {a.train_meshes} meshes x {a.train_windows} windows ({ts.dm.N} cells) in {gen_s:.1f}s")
    t1 = time.time()
    hist = lc.train(data[:-4], a.train_iters, log=log)
    torch.cuda.synchronize()
    train_s = time.time() - t1
    held = lc.eval_loss(data[-4:])
    zero = float(np.mean([float((T.double() ** 2).mean()) for _, T in data[-4:]]))
    log(f"trained {a.train_iters} iters in {train_s:.1f}s; held-out loss {held:.4f} (base predictor alone {zero:.4f})")
    stats, weights = lc.stats, lc.net.state_dict()
    del data, lc, ts
    torch.cuda.empty_cache()

    ep, eparams = random_channel_batch(a.eval_meshes, seed=a.seed + 2000, n_min=a.n_min, n_max=a.n_max, device=dev)
    plain = BatchedSolver(ep, device=dev)
    dts = plain.cfl_dt(0.5)
    plain.set_dt(dts)
    learn = BatchedSolver(ep, device=dev)
    learn.set_dt(dts)
    lc = LearnedCorrector(learn, k=a.window, layers=a.layers, tol=a.tol, settle=a.settle)
    lc.stats = stats
    lc._apply_stats()
    lc.net.load_state_dict(weights)
    g = plain.capture_graph(a.window)                 # one window of plain steps as a HIP graph
    for _ in range(a.eval_spinup):                    # past the impulsive start
        g.replay()
    torch.cuda.synchronize()
    U0 = plain.state.clone()
    # plain timing: graph replays only (the captured steps have the solver's buffer
    # ping-pong baked in, so no eager step may come between replays)
    g.replay()
    torch.cuda.synchronize()
    tp0 = time.perf_counter()
    for _ in range(a.eval_windows):
        g.replay()
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - tp0

    def counts():
        return np.array([lc.n_accept, lc.n_partial, lc.n_reject, lc.mesh_accepted, lc.mesh_windows])

    # what the network gains per window: predictions from states of the plain trajectory
    # (history from plain steps) against the plain RK4 window, network and base alone
    def mesh_rel(X, R):
        out = []
        for i in range(a.eval_meshes):
            lo, hi = int(plain.offsets[i]), int(plain.offsets[i + 1])
            out.append(float((X[lo:hi] - R[lo:hi]).double().norm() / R[lo:hi].double().norm()))
        return np.asarray(out)

    learn.set_state(U0)
    lc.plain_window()
    werr = {"net": [], "base": []}
    U_end = torch.empty_like(U0)
    for _ in range(4):
        U1n = lc.predict(True)[1].clone()
        U1b = lc.predict(False)[1].clone()
        lc.plain_window(end_out=U_end)
        werr["net"].append(mesh_rel(U1n, U_end))
        werr["base"].append(mesh_rel(U1b, U_end))
    werr = {k: np.concatenate(v) for k, v in werr.items()}
    log(f"one window from the plain trajectory, end-state rel-L2 vs plain RK4 (64 meshes x 4 windows): network "
        f"median {np.median(werr['net']):.2e} max {werr['net'].max():.2e}; base alone median "
        f"{np.median(werr['base']):.2e} max {werr['base'].max():.2e}")

    learn.set_state(U0)                               # back to the timed runs' start
    lc.plain_window()
    c0 = counts()
    learn_s, acc = lc.run(a.eval_windows)
    c_learn = counts() - c0
    U_learn = learn.state.clone()
    # the method without the network: the same windows with the base predictor alone
    learn.set_state(U0)
    lc.plain_window()
    c0 = counts()
    base_s, acc_base = lc.run(a.eval_windows, use_net="base")
    c_base = counts() - c0
    U_base = learn.state.clone()
    sim_t = a.eval_windows * (a.window + a.settle) * dts      # per mesh, corrector runs
    sim_p = a.eval_windows * a.window * dts                   # per mesh, the timed plain graph
    # the plain state at the corrector runs' time: (windows + 1) (k + settle) eager steps
    # from U0 (the corrector's first window is plain, for its history)
    plain.set_state(U0)
    for _ in range((a.eval_windows + 1) * (a.window + a.settle)):
        plain.step()
    torch.cuda.synchronize()
    N = plain.dm.N
    log(f"eval: {a.eval_meshes} meshes, {N} cells, {a.eval_windows} windows of {a.window} steps after "
        f"{a.eval_spinup + 1} spin-up windows: plain {plain_s:.3f}s, learned {learn_s:.3f}s (windows accepted/"
        f"partial/rejected {c_learn[:3]}, mesh predictions accepted {c_learn[3]}/{c_learn[4]}), base only "
        f"{base_s:.3f}s ({c_base[:3]}, {c_base[3]}/{c_base[4]})")

    vs_plain = {"learned": mesh_rel(U_learn, plain.state), "base": mesh_rel(U_base, plain.state)}
    log(f"vs plain fp32 RK4 (all meshes): learned max {vs_plain['learned'].max():.3e} "
        f"median {np.median(vs_plain['learned']):.3e}; base max {vs_plain['base'].max():.3e}")

    # CPU leg (the checker): the FP64 CPU oracle re-runs the smallest evaluation meshes
    errs = {}
    order = [] if a.no_cpu_baseline else list(np.argsort([p["cells"] for p in eparams])[:a.oracle_meshes])
    if len(order) and int(np.argmax(vs_plain["learned"])) not in order:
        order.append(int(np.argmax(vs_plain["learned"])))
    if len(order):
        from oracle import fvm_oracle as O
    for i in order:
        pts, cells, fi, fb, tags = eparams[i]["mesh_tuple"]
        cfg = ep[i][2]
        _, ofvm = O.from_mesh_tuple(pts, cells, fi, fb, tags, _oracle_phys(cfg), dtype=torch.float64)
        st = O.OracleStepper(ofvm, "Butcher", "RK4", dt=float(dts[i]))
        lo, hi = int(plain.offsets[i]), int(plain.offsets[i + 1])
        U = U0[lo:hi].cpu().double()
        for _ in range((a.eval_windows + 1) * (a.window + a.settle)):
            U = st.step(U)
        ref = U.numpy()
        rel = lambda x: (np.linalg.norm(x.cpu().double().numpy() - ref, axis=0) / np.linalg.norm(ref, axis=0)).tolist()
        full = lambda x: float(np.linalg.norm(x.cpu().double().numpy() - ref) / np.linalg.norm(ref))
        errs[int(i)] = {"cells": int(eparams[i]["cells"]), "learned": full(U_learn[lo:hi]),
                        "base": full(U_base[lo:hi]), "plain": full(plain.state[lo:hi]),
                        "learned_per_variable": rel(U_learn[lo:hi]), "plain_per_variable": rel(plain.state[lo:hi])}
        log(f"mesh {i} ({eparams[i]['cells']} cells): state rel-L2 vs FP64 learned {errs[int(i)]['learned']:.3e} "
            f"(per variable {errs[int(i)]['learned_per_variable']}), base {errs[int(i)]['base']:.3e}, "
            f"plain {errs[int(i)]['plain']:.3e}")
    worst = max(e["learned"] for e in errs.values()) if errs else None
    out = {
        "metric": "simulated time per wall second, learned window corrector vs plain RK4 (batched 2-D channel)",
        "value": float(np.mean(sim_t) / learn_s),
        "unit": "sim-s/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "vs_plain": float(np.mean(sim_t) / learn_s) / float(np.mean(sim_p) / plain_s),
        "plain_value": float(np.mean(sim_p) / plain_s),
        "accepted_windows": int(acc),
        "partial_windows": int(c_learn[1]),
        "rejected_windows": int(c_learn[2]),
        "mesh_predictions_accepted": f"{int(c_learn[3])}/{int(c_learn[4])}",
        "windows": a.eval_windows,
        "base_only": {"value": float(np.mean(sim_t) / base_s), "accepted_windows": int(acc_base),
                      "partial_windows": int(c_base[1]),
                      "mesh_predictions_accepted": f"{int(c_base[3])}/{int(c_base[4])}",
                      "vs_plain": float(np.mean(sim_t) / base_s) / float(np.mean(sim_p) / plain_s)},
        "window_error_vs_rk4": {"network_median": float(np.median(werr["net"])),
                                "network_max": float(werr["net"].max()),
                                "base_median": float(np.median(werr["base"])), "base_max": float(werr["base"].max()),
                                "note": "end state of one predicted window vs the plain RK4 window from the same "
                                        "state of the plain trajectory, per mesh, 4 windows"},
        "max_state_rel_l2_vs_plain_fp32": {"learned": float(vs_plain["learned"].max()),
                                           "base": float(vs_plain["base"].max())},
        "max_state_rel_l2_vs_fp64": worst,
        "error_norm": "relative L2 over each mesh's whole conserved state (per-variable values in config)",
        "tol": a.tol,
        "dtype": "bf16 network (fp32 accumulation), f32 solver",
        "data": "synthetic: trajectories of the batched solver on seeded random channel meshes",
        "config": {"workload": f"learned corrector, {a.layers} message-passing layers, hidden 128, window of "
                               f"{a.window} RK4 steps at CFL 0.5 + {a.settle} plain settle step(s)",
                   "train_meshes": a.train_meshes,
                   "train_pairs": a.train_meshes * a.train_windows, "train_iters": a.train_iters,
                   "eval_meshes": a.eval_meshes, "eval_cells": N, "cells_range": [a.n_min, a.n_max],
                   "eval_spinup_windows": a.eval_spinup,
                   "held_out_loss": held, "base_window_loss": zero, "train_seconds": train_s,
                   "loss_history": hist[-3:], "errors_vs_fp64": errs},
    }
    print(json.dumps(out), flush=True)


def corrector_timing_main(a):
    """Stream-event medians of the pieces of one corrector window (scales, network input,
    network forward, predicted states, the two check RHS, the defect reduction), of the
    settle step, of a partly re-run window (tile-masked plain RK4 of a share of the meshes,
    f at the merged midpoint, the settle step) at several shares, and of a plain RK4 window
    (HIP graph), on the evaluation batch of --learned."""
    from fvm_solver_amd import _native as nat
    from fvm_solver_amd.batch import BatchedSolver, random_channel_batch
    from fvm_solver_amd.learned.corrector import LearnedCorrector
    dev = torch.device("cuda:0")
    ep, _ = random_channel_batch(a.eval_meshes, seed=a.seed + 2000, n_min=a.n_min, n_max=a.n_max, device=dev)
    s = BatchedSolver(ep, device=dev)
    s.set_dt(s.cfl_dt(0.5))
    lc = LearnedCorrector(s, k=a.window, layers=a.layers, tol=a.tol, settle=a.settle)
    lc.set_stats(s.state)
    lc.plain_window()
    torch.cuda.synchronize()
    W = lc.window()
    names = ["scales", "features", "network", "predict", "check_rhs"]
    res = {n: [] for n in names}
    (am, bm), (ae, be) = lc.base_mid, lc.base_end
    for it in range(8):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
        e[0].record()
        lc._scales(lc._f0, lc._fh, W, lc._s1, lc._s2, lc._st)
        e[1].record()
        X = lc._features_from(lc._prim0, lc._f0, lc._fh, W, lc._s1, lc._s2, lc._X)
        e[2].record()
        y = lc.net.forward(X, lc.nbr)
        e[3].record()
        nat.check(nat.lib().fvm_mp_window_predict(nat.ptr(s.state), nat.ptr(lc._f0), nat.ptr(lc._fh), am, bm, ae, be,
                                                  nat.ptr(W), nat.ptr(y), 2 * lc.C, nat.ptr(lc._st),
                                                  nat.ptr(lc.cell_mesh), lc.N, lc.C, nat.ptr(lc._Um),
                                                  nat.ptr(lc._U1), nat.stream_ptr(dev)), "fvm_mp_window_predict")
        e[4].record()
        lc._rhs(lc._Um, lc._prim1, lc._fm)
        lc._rhs(lc._U1, lc._prim1, lc._f1)
        e[5].record()
        torch.cuda.synchronize()
        if it >= 2:
            for i, n in enumerate(names):
                res[n].append(e[i].elapsed_time(e[i + 1]))
    def timed(fn, reps=5):
        ts = []
        for it in range(reps + 1):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            if it:
                ts.append(e0.elapsed_time(e1))
        return float(np.median(ts))

    def defect():
        nat.check(nat.lib().fvm_mp_defect3(nat.ptr(s.state), nat.ptr(lc._Um), nat.ptr(lc._U1), nat.ptr(lc._f0),
                                           nat.ptr(lc._fm), nat.ptr(lc._f1), nat.ptr(W), nat.ptr(lc.off), lc.M, lc.C,
                                           nat.ptr(lc._def), nat.ptr(lc._part), nat.stream_ptr(dev)), "fvm_mp_defect3")

    res["defect"] = [timed(defect)]
    settle_ms = timed(lc._settle_and_start)
    rng = np.random.default_rng(a.seed)
    partial = {}
    for share in (0.25, 0.5, 0.75, 1.0):
        redo = rng.random(lc.M) < share
        redo[int(rng.integers(lc.M))] = True
        cells = int(sum(int(s.offsets[i + 1] - s.offsets[i]) for i in np.flatnonzero(redo)))
        partial[str(share)] = {"meshes": int(redo.sum()), "cells": cells,
                               "ms": timed(lambda: lc._partial_window(~redo), 3)}
    g = s.capture_graph(a.window)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        g.replay()
    e1.record()
    torch.cuda.synchronize()
    out = {"cells": int(s.dm.N), "meshes": a.eval_meshes, "ms": {n: float(np.median(v)) for n, v in res.items()},
           "settle_and_next_f_ms": settle_ms, "partial_window_ms": partial,
           "plain_window_ms": e0.elapsed_time(e1) / 5}
    out["corrector_window_ms"] = float(sum(out["ms"].values()))
    print(json.dumps(out), flush=True)


def cpu_config1_main(a):
    """The 20k-cell 2-D channel, 1000 RK4 steps at CFL 0.5, on the CPU restatement of
    the reference (no GPU is used): the reference CPU path's own configuration."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import fvm_oracle as O
    from fvm_solver_amd.config_fvm import ConfigChannel
    from fvm_solver_amd.mesh_gen.channel import channel_mesh_2d
    n_cells = a.cells or 20_000
    pts, cells, (fi, fb), tags = channel_mesh_2d(n_cells, seed=a.seed)
    cfg = ConfigChannel(device="cpu", n_cells=n_cells, solver_name="Butcher", solver_extra="RK4")
    _, ofvm = O.from_mesh_tuple(pts, cells, fi, fb, tags, _oracle_phys(cfg))
    N, D = len(cells), 2
    c = pts[cells].mean(1)
    prim = torch.zeros(N, 4, dtype=torch.float32)
    prim[:, 0] = cfg.inlet_cfg.v_n_inf
    r2 = torch.from_numpy(((c - np.asarray(cfg.bump_centre)[:D]) ** 2).sum(1)).float()
    prim[:, 2] = 1.0 + cfg.bump_amp * torch.exp(-r2 / (2 * cfg.bump_sigma ** 2))
    prim[:, 3] = cfg.inlet_cfg.T_inf
    U0 = torch.cat(ofvm.primatives_to_state(prim[:, :2], prim[:, 2:3], prim[:, 3:]), -1)
    dt = float(np.float32(ofvm.cfl_dt(prim.double(), cfg.cfl)))
    cpu = host_cpu()
    threads = min(x for x in (cpu["physical_cores"], cpu["affinity_cpus"], cpu["logical_cpus"]) if x)
    res = {}
    for nt, steps in ((threads, 1000), (1, 100)):
        torch.set_num_threads(nt)
        st = O.OracleStepper(ofvm, "Butcher", "RK4", dt=dt)
        U = st.step(U0.clone())
        t0 = time.perf_counter()
        for i in range(steps):
            U = st.step(U)
            if i % 200 == 199:
                log(f"{nt} threads: step {i + 1}")
        res[nt] = (steps, time.perf_counter() - t0)
    out = {"metric": "cell-updates/s (CPU restatement of the reference, 2-D NS channel, RK4)",
           "value": N * res[threads][0] / res[threads][1], "unit": "cell-updates/s", "higher_is_better": True,
           "config": {"workload": f"2-D channel, {N} cells, RK4 at CFL 0.5, dt {dt:.4e}", "steps": res[threads][0],
                      "seconds": res[threads][1], "threads": threads,
                      "single_thread": {"steps": res[1][0], "seconds": res[1][1],
                                        "value": N * res[1][0] / res[1][1]},
                      "cpu_model": cpu["model"], "host_physical_cores": cpu["physical_cores"],
                      "affinity_cpus": cpu["affinity_cpus"]}}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    _a = _parse()
    if _a.cpu_config1:
        cpu_config1_main(_a)
    elif _a.corrector_timing:
        corrector_timing_main(_a)
    elif _a.learned:
        if _a.gpus != 1:
            print("bench.py runs on one GPU only", file=sys.stderr)
            sys.exit(2)
        learned_main(_a)
    elif _a.config == "batch":
        if _a.gpus != 1:
            print("bench.py runs on one GPU only", file=sys.stderr)
            sys.exit(2)
        batch_main(_a)
    else:
        main()
