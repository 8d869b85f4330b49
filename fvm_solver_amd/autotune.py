"""Launch-shape autotuning of the per-cell kernels.

The tunable is the workgroup count of k_grad / k_facet_flux / k_div (``FvmMeshDev.persist_blocks``):
0 launches one workgroup per 256-cell tile; a positive multiple of 8 launches that many
workgroups which walk the tiles grid-stride (persistent), each XCD over a contiguous
tile range.  Candidates are timed with stream events on the live problem (the steps it
runs are real steps of the simulation) and the fastest is kept; results are cached per
(cells, dim).
"""
from __future__ import annotations

import torch

_CACHE: dict = {}


def candidates(n_cells: int, n_cu: int | None = None):
    if n_cu is None:
        n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    tiles = (n_cells + 255) // 256
    out = [0]
    for per_cu in (2, 4, 8, 16):
        g = per_cu * n_cu
        g -= g % 8
        if 0 < g < tiles:
            out.append(g)
    return out


def tune(step, device_mesh, n_cells: int, dim: int, reps: int = 4, verbose: bool = False) -> dict:
    """Time ``step()`` (one full time step) for each candidate; keep the fastest."""
    key = (n_cells, dim)
    if key in _CACHE:
        device_mesh.set_persist_blocks(_CACHE[key]["persist_blocks"])
        return _CACHE[key]
    results = {}
    for g in candidates(n_cells):
        device_mesh.set_persist_blocks(g)
        step()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            step()
        e1.record()
        torch.cuda.synchronize()
        results[g] = e0.elapsed_time(e1) / reps
        if verbose:
            print(f"[autotune] persist_blocks={g}: {results[g] * 1e3:.1f} us/step", flush=True)
    best = min(results, key=results.get)
    device_mesh.set_persist_blocks(best)
    _CACHE[key] = {"persist_blocks": best, "us_per_step": {k: v * 1e3 for k, v in results.items()}}
    return _CACHE[key]
