"""ELL SpMV/SpMM and CSR->ELL packing on gfx950 -- drop-in for ``time_fvm/utils/ell_kernel.py``.

Same call signatures and shape rules as the reference (ell_kernel.py:135-256):
``ell_spmv(vals, cols, x, b=None)`` keeps x's rank, ``ell_spmm`` dispatches B == 1 to
``ell_spmv``, ``csr_to_ell(csr, K=None)`` returns ``(vals, cols)`` of shape [M, K] and
raises ValueError when a row holds more than K entries.  The products run in the HIP
kernels of libfvm_hip.so (one row per lane, accumulation in slot order); vals/cols may
be in any stride layout and cols may be int32 or int64.

``csr_to_ell`` packs slot-major (value k of row r at k * M + r, so a wavefront's slot
load is one contiguous segment) and returns [M, K] views of that storage with int32
column indices; the index values are those of the reference.
"""
from __future__ import annotations

import torch

from fvm_solver_amd import _native as nat


def _spmm_into(vals: torch.Tensor, cols: torch.Tensor, x2: torch.Tensor, b, y: torch.Tensor):
    M, K = vals.shape
    N, B = x2.shape
    for t, n in ((vals, "vals"), (cols, "cols"), (x2, "x"), (y, "y")):
        nat.require_device(t, n)
    if vals.dtype != torch.float32 or x2.dtype != torch.float32:
        raise TypeError("ell kernels compute in float32 (vals and x must be float32)")
    if cols.dtype not in (torch.int32, torch.int64):
        raise TypeError("cols must be int32 or int64")
    if cols.shape != vals.shape:
        raise ValueError(f"vals {tuple(vals.shape)} and cols {tuple(cols.shape)} differ in shape")
    if b is not None:
        b = b.reshape(-1)
        if b.numel() != M:
            raise ValueError(f"bias has {b.numel()} entries, expected {M}")
        nat.require_device(b, "b")
        b = b.contiguous()
    rc = nat.lib().fvm_ell_spmm_f32(
        nat.ptr(vals), vals.stride(0), vals.stride(1), nat.ptr(cols), int(cols.dtype == torch.int64),
        cols.stride(0), cols.stride(1), M, K, nat.ptr(x2), x2.stride(0), x2.stride(1), B, nat.ptr(b),
        nat.ptr(y), y.stride(0), y.stride(1), nat.stream_ptr(x2.device))
    nat.check(rc, "fvm_ell_spmm_f32")
    return y


def ell_spmv(vals, cols, x, b=None):
    """y = A_ell @ x (+ b); x is [N] or [N, 1], y matches x's rank (ell_kernel.py:135-173)."""
    M, K = vals.shape
    x_1d = x.ndim == 1
    x2 = x.unsqueeze(-1) if x_1d else x
    N, B = x2.shape
    y = torch.empty((M, B), device=x.device, dtype=x.dtype)
    _spmm_into(vals, cols, x2, b, y)
    return y.squeeze(-1) if x_1d else y


def ell_spmm(vals, cols, x, b=None):
    """Y = A_ell @ X (+ b[:, None]); X is [N, B] (ell_kernel.py:176-210)."""
    M, K = vals.shape
    N, B = x.shape
    if B == 1:
        return ell_spmv(vals, cols, x, b=b)
    y = torch.empty((M, B), device=x.device, dtype=x.dtype)
    return _spmm_into(vals, cols, x, b, y)


def print_best_config():
    """The reference prints Triton's autotuned block shape (ell_kernel.py:213-222); these
    kernels have one fixed launch shape: 256-lane blocks, one row per lane."""
    print("HIP ELL kernel: 256 threads/block, one row per lane, slot-major ELL")


def csr_to_ell_slot_major(csr: torch.Tensor, K: int | None = None):
    """Pack a sparse CSR matrix into slot-major ELL storage: (vals [K, M] f32, cols [K, M] i32)."""
    if csr.layout != torch.sparse_csr:
        raise TypeError("Input must be a sparse CSR tensor")
    dev = csr.device
    nat.require_device(torch.empty(0, device=dev), "csr")
    crow = csr.crow_indices().to(torch.int64).contiguous()
    col = csr.col_indices().to(torch.int64).contiguous()
    val = csr.values().to(torch.float32).contiguous()
    M = csr.shape[0]
    s = nat.stream_ptr(dev)
    if K is None:
        kmax = torch.zeros(1, dtype=torch.int32, device=dev)
        nat.check(nat.lib().fvm_csr_max_row_nnz(nat.ptr(crow), M, nat.ptr(kmax), s), "fvm_csr_max_row_nnz")
        K = int(kmax.item())
    ev = torch.empty((K, M), dtype=torch.float32, device=dev)
    ec = torch.empty((K, M), dtype=torch.int32, device=dev)
    flag = torch.zeros(1, dtype=torch.int32, device=dev)
    nat.check(nat.lib().fvm_ell_pack_f32(nat.ptr(crow), nat.ptr(col), nat.ptr(val), M, int(K), nat.ptr(ev),
                                         nat.ptr(ec), M, nat.ptr(flag), s), "fvm_ell_pack_f32")
    if int(flag.item()):
        raise ValueError("Some rows have more nonzeros than K")
    return ev, ec


def csr_to_ell(csr: torch.Tensor, K: int | None = None):
    """CSR -> ELL (ell_kernel.py:225-256).  Returns [M, K] views (vals f32, cols int32)."""
    ev, ec = csr_to_ell_slot_major(csr, K)
    return ev.t(), ec.t()
