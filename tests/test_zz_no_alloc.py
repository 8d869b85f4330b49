"""The allocation-free conservation entry point (fvm_conservation_f64_noalloc, whose
partials live in the library's own device scratch array) against the entry point that
takes a caller workspace (fvm_conservation_f64): the same kernels in the same order, so
the fp64 totals must be bitwise equal -- on two calls in a row, the second with a
smaller input (the scratch keeps the first call's partials, which must not leak into
the second) -- and a variable count beyond the scratch limit is rejected.  Both are
called through the C ABI."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _lib():
    from fvm_solver_amd import _native as nat
    return nat, nat.lib()


def _with_workspace(nat, lib, U, vol):
    N, C = U.shape
    out = torch.full((C,), float("nan"), dtype=torch.float64, device=DEV)
    part = torch.empty(256 * C, dtype=torch.float64, device=DEV)
    rc = lib.fvm_conservation_f64(nat.ptr(U), nat.ptr(vol), N, C, nat.ptr(out), nat.ptr(part),
                                  nat.stream_ptr(torch.device(DEV)))
    nat.check(rc, "fvm_conservation_f64")
    return out


def _no_alloc(nat, lib, U, vol):
    N, C = U.shape
    out = torch.full((C,), float("nan"), dtype=torch.float64, device=DEV)
    rc = lib.fvm_conservation_f64_noalloc(nat.ptr(U), nat.ptr(vol), N, C, nat.ptr(out),
                                          nat.stream_ptr(torch.device(DEV)))
    nat.check(rc, "fvm_conservation_f64_noalloc")
    return out


def test_no_alloc_matches_workspace_entry_point():
    nat, lib = _lib()
    g = torch.Generator().manual_seed(3)
    for N, C in ((1_000_003, 5), (70_001, 4)):          # second call smaller (fewer cells and variables)
        print(f"[gpu-test] conservation no-alloc vs workspace, N={N} C={C}", flush=True)
        U = torch.randn(N, C, generator=g).to(DEV)
        vol = torch.rand(N, generator=g, dtype=torch.float64).to(DEV)
        a = _no_alloc(nat, lib, U, vol)
        b = _with_workspace(nat, lib, U, vol)
        torch.cuda.synchronize()
        assert torch.equal(a, b), (a, b)
        ref = (vol.cpu()[:, None] * U.cpu().double()).sum(0)
        assert torch.allclose(a.cpu(), ref, rtol=1e-12, atol=1e-7)


def test_no_alloc_rejects_too_many_variables():
    nat, lib = _lib()
    U = torch.zeros(16, 9, device=DEV)
    vol = torch.ones(16, dtype=torch.float64, device=DEV)
    out = torch.zeros(9, dtype=torch.float64, device=DEV)
    rc = lib.fvm_conservation_f64_noalloc(nat.ptr(U), nat.ptr(vol), 16, 9, nat.ptr(out),
                                          nat.stream_ptr(torch.device(DEV)))
    assert rc == 1
    assert b"scratch limit" in lib.fvm_last_error()
