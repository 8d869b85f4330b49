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


# Launch shapes of the fast ELL kernel the autotuner tries: (rows per lane, threads per
# block).  All compute bitwise the same products.
SHAPES = tuple((r, t) for r in (1, 2, 4) for t in (64, 128, 256, 512))
_TUNED: dict = {}          # (M, B, K, has_bias) -> {"shape": (rows_per_lane, block), "us": {...}}
_TUNE_REPS = 5


def _launch(args, shape, stream):
    return nat.lib().fvm_ell_spmm_f32_shape(*args, shape[0], shape[1], stream)


def _autotune(key, args, stream):
    """Time every launch shape on the live operands (the product is written each time,
    identically) and keep the fastest for ``key``, as the reference's Triton autotune does
    on the first call per key (utils/ell_kernel.py:8-23, 83-98)."""
    us = {}
    for shape in SHAPES:
        nat.check(_launch(args, shape, stream), "fvm_ell_spmm_f32_shape")      # warm-up
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(_TUNE_REPS):
            _launch(args, shape, stream)
        e1.record()
        torch.cuda.synchronize()
        us[shape] = 1e3 * e0.elapsed_time(e1) / _TUNE_REPS
    best = min(us, key=us.get)
    _TUNED[key] = {"shape": best, "us": {f"{r}x{t}": v for (r, t), v in us.items()}}
    return best


def best_config(M: int, B: int, K: int, has_bias: bool = False):
    """The tuned launch shape for a key, or None before its first product."""
    return _TUNED.get((M, B, K, bool(has_bias)))


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
    is64 = int(cols.dtype == torch.int64)
    args = (nat.ptr(vals), vals.stride(0), vals.stride(1), nat.ptr(cols), is64, cols.stride(0), cols.stride(1), M, K,
            nat.ptr(x2), x2.stride(0), x2.stride(1), B, nat.ptr(b), nat.ptr(y), y.stride(0), y.stride(1))
    stream = nat.stream_ptr(x2.device)
    shape = (1, 256)
    if M > 0 and nat.lib().fvm_ell_fast_path(vals.stride(0), vals.stride(1), is64, cols.stride(0), cols.stride(1), K,
                                              x2.stride(0), x2.stride(1), B, y.stride(0), y.stride(1),
                                              nat.ptr(x2), nat.ptr(y)):
        key = (M, B, K, b is not None)
        tuned = _TUNED.get(key)
        shape = tuned["shape"] if tuned is not None else _autotune(key, args, stream)
    nat.check(_launch(args, shape, stream), "fvm_ell_spmm_f32")
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
    """Print the launch shapes the autotuner chose, per (M, B, K, has_bias) key, as the
    reference prints its Triton best config (ell_kernel.py:213-222)."""
    if not _TUNED:
        print("No best config yet — run the kernel first.")
        return
    for (M, B, K, hb), t in _TUNED.items():
        r, th = t["shape"]
        print(f"HIP ELL config M={M} B={B} K={K} bias={hb}: {r} row(s) per lane, {th} threads per block "
              f"({t['us'][f'{r}x{th}']:.1f} us)")


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
