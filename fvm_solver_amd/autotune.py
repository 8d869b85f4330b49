"""Launch-shape and pipeline autotuning of the RK stage.

Candidates:
  * ``"fused"`` -- the tile-fused stage (one workgroup per 256-cell tile, one launch per
    RHS), when the mesh has tile tables;
  * ``g`` (an int) -- the three-kernel pipeline (k_grad / k_facet_flux / k_div) with
    ``FvmMeshDev.persist_blocks = g``: 0 launches one workgroup per 256-cell tile; a
    positive multiple of 8 launches that many workgroups which walk the tiles
    grid-stride (persistent), each XCD over a contiguous tile range.
Candidates are timed with stream events on the live problem (the steps it runs are real
steps of the simulation; all candidates compute bitwise the same states) and the fastest
is kept; results are cached per (cells, dim).
"""
from __future__ import annotations

import torch

_CACHE: dict = {}


def candidates(n_cells: int, n_cu: int | None = None, fused: bool = False):
    if n_cu is None:
        n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    tiles = (n_cells + 255) // 256
    out = ["fused"] if fused else []
    out.append(0)
    for per_cu in (2, 4, 8, 16):
        g = per_cu * n_cu
        g -= g % 8
        if 0 < g < tiles:
            out.append(g)
    return out


def _apply(E, c):
    if c == "fused":
        E.set_fused(True)
    else:
        E.set_fused(False)
        E.mesh.set_persist_blocks(c)


def tune(step, E, n_cells: int, dim: int, reps: int = 4, verbose: bool = False) -> dict:
    """Time ``step()`` (one full time step) for each candidate on FacetCalc ``E``; keep
    the fastest.  Returns {"choice", "persist_blocks", "fused", "us_per_step"}."""
    key = (n_cells, dim, E.mesh.tiles is not None)
    if key in _CACHE:
        _apply(E, _CACHE[key]["choice"])
        return _CACHE[key]
    results = {}
    for c in candidates(n_cells, fused=E.mesh.tiles is not None):
        _apply(E, c)
        step()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            step()
        e1.record()
        torch.cuda.synchronize()
        results[c] = e0.elapsed_time(e1) / reps
        if verbose:
            print(f"[autotune] {c}: {results[c] * 1e3:.1f} us/step", flush=True)
    best = min(results, key=results.get)
    _apply(E, best)
    _CACHE[key] = {"choice": best, "fused": best == "fused",
                   "persist_blocks": 0 if best == "fused" else best,
                   "us_per_step": {str(k): v * 1e3 for k, v in results.items()}}
    return _CACHE[key]
