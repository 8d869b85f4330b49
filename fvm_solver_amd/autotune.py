"""Launch-shape and pipeline autotuning of the RK stage.

Candidates:
  * ``"fused"`` / ``"fused320"`` / ``"fused384"`` / ``"fused512"`` -- the tile-fused
    stage (one workgroup per 256-cell tile, one launch per RHS) with 256 to 512 threads
    per tile (those the mesh's tile tables allow), when the mesh has tile tables;
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


def _fused_name(nt: int) -> str:
    return "fused" if nt == 256 else f"fused{nt}"


def candidates(n_cells: int, n_cu: int | None = None, fused_threads=()):
    if n_cu is None:
        n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    tiles = (n_cells + 255) // 256
    out = [_fused_name(nt) for nt in fused_threads]
    out.append(0)
    for per_cu in (2, 4, 8, 16):
        g = per_cu * n_cu
        g -= g % 8
        if 0 < g < tiles:
            out.append(g)
    return out


def _threads_of(c) -> int:
    return 256 if c == "fused" else int(c[5:])


def _apply(E, c):
    if isinstance(c, str):
        E.mesh.set_fused_threads(_threads_of(c))
        E.set_fused(True)
    else:
        E.set_fused(False)
        E.mesh.set_persist_blocks(c)


def tune(step, E, n_cells: int, dim: int, reps: int = 4, rounds: int = 2, verbose: bool = False) -> dict:
    """Time ``step()`` (one full time step) for each candidate on FacetCalc ``E``; keep
    the fastest.  Each candidate is timed over ``reps`` steps in each of ``rounds``
    rounds (candidates interleaved, the best round kept, so a slow moment of the device
    does not decide).  Returns {"choice", "persist_blocks", "fused", "fused_threads",
    "us_per_step"}."""
    key = (n_cells, dim, E.mesh.tiles is not None)
    if key in _CACHE:
        _apply(E, _CACHE[key]["choice"])
        return _CACHE[key]
    results = {}
    cands = candidates(n_cells, fused_threads=E.mesh.fused_thread_options())
    for _ in range(rounds):
        for c in cands:
            _apply(E, c)
            step()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                step()
            e1.record()
            torch.cuda.synchronize()
            t = e0.elapsed_time(e1) / reps
            results[c] = min(t, results.get(c, t))
            if verbose:
                print(f"[autotune] {c}: {t * 1e3:.1f} us/step", flush=True)
    best = min(results, key=results.get)
    _apply(E, best)
    fused = isinstance(best, str)
    _CACHE[key] = {"choice": best, "fused": fused, "fused_threads": _threads_of(best) if fused else 0,
                   "persist_blocks": 0 if fused else best,
                   "us_per_step": {str(k): v * 1e3 for k, v in results.items()}}
    return _CACHE[key]
