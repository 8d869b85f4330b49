"""Sparse operator API on gfx950 -- drop-in for ``time_fvm/utils/sparse.py``.

``SPM`` keeps the reference's interface (sparse.py:7-21): ``spMV(x, b=None)`` and
``spMM(X, b=None)`` compute A x + b.  ``to_sparse(A, device)`` (sparse.py:65-70) returns
``SPMHip`` for a GPU device; the operator is packed once into slot-major ELL on the
device and every product runs in libfvm_hip.so.  There is no CPU implementation here:
a non-GPU device raises.  The COO builders (sparse.py:73-325) are provided with the same
semantics for code that assembles operators through them.
"""
from __future__ import annotations

from abc import ABC, abstractmethod

import torch

from fvm_solver_amd.utils.ell_kernel import _spmm_into, csr_to_ell_slot_major


class SPM(ABC):
    """Sparse matrix (sparse.py:7-21)."""

    def __init__(self, A, device):
        self.shape = A.shape
        self.device = device

    @abstractmethod
    def spMV(self, x: torch.Tensor, b: torch.Tensor = None) -> torch.Tensor:
        raise NotImplementedError("SPM is an abstract base class")

    @abstractmethod
    def spMM(self, X: torch.Tensor, b: torch.Tensor = None) -> torch.Tensor:
        raise NotImplementedError("SPM is an abstract base class")


class SPMHip(SPM):
    """Slot-major ELL operator on the GPU (replaces SPMCuda, sparse.py:24-38)."""

    def __init__(self, A: torch.Tensor, device):
        super().__init__(A, device)
        if A.layout != torch.sparse_csr:
            A = A.coalesce().to_sparse_csr() if A.layout == torch.sparse_coo else A.to_sparse_csr()
        A = A.to(device)
        self.ell_vals, self.ell_cols = csr_to_ell_slot_major(A)      # [K, M]
        self.vals, self.cols = self.ell_vals.t(), self.ell_cols.t()   # [M, K] views

    def spMM(self, X: torch.Tensor, b: torch.Tensor = None) -> torch.Tensor:
        M = self.shape[0]
        if X.ndim != 2:
            raise ValueError("spMM expects X of shape [n, B]")
        y = torch.empty((M, X.shape[1]), device=X.device, dtype=X.dtype)
        return _spmm_into(self.vals, self.cols, X, b, y)

    def spMV(self, x: torch.Tensor, b: torch.Tensor = None) -> torch.Tensor:
        M = self.shape[0]
        x2 = x.unsqueeze(-1) if x.ndim == 1 else x
        y = torch.empty((M, x2.shape[1]), device=x.device, dtype=x.dtype)
        _spmm_into(self.vals, self.cols, x2, b, y)
        return y.squeeze(-1) if x.ndim == 1 else y


SPMCuda = SPMHip


def to_sparse(A: torch.Tensor, device) -> SPM:
    """Operator factory (sparse.py:65-70).  GPU devices only."""
    device = torch.device(device)
    if device.type != "cuda":
        raise RuntimeError(f"to_sparse: device {device} is not a GPU; this library only runs the HIP path")
    return SPMHip(A, device)


def lift_sparse_matrix(A_old: torch.Tensor, n_comp: int):
    """(M, N) COO -> (M*n_comp, N*n_comp) block-diagonal lift, row i*n_comp+c (sparse.py:231-278)."""
    A_old = A_old.coalesce()
    idx, v = A_old.indices(), A_old.values()
    M, N = A_old.size()
    comp = torch.arange(n_comp, device=idx.device)
    rows = (idx[0][:, None] * n_comp + comp[None]).reshape(-1)
    cols = (idx[1][:, None] * n_comp + comp[None]).reshape(-1)
    vals = v[:, None].expand(-1, n_comp).reshape(-1)
    return torch.sparse_coo_tensor(torch.stack([rows, cols]), vals, (M * n_comp, N * n_comp)).coalesce()


def interleave_sparse_rows(tensors: list[torch.Tensor]) -> torch.Tensor:
    """Interleave k (m, n) COO matrices row-wise into (k*m, n): row k*r + i (sparse.py:281-325)."""
    if not tensors:
        raise ValueError("Must provide at least one tensor")
    m, n = tensors[0].shape
    for t in tensors:
        if t.layout != torch.sparse_coo:
            raise TypeError("All tensors must be sparse COO tensors")
        if t.ndim != 2 or tuple(t.shape) != (m, n):
            raise ValueError("All tensors must have shape (m, n)")
    k = len(tensors)
    idx, val = [], []
    for i, t in enumerate(tensors):
        t = t.coalesce()
        ii = t.indices().clone()
        ii[0] = k * ii[0] + i
        idx.append(ii)
        val.append(t.values())
    return torch.sparse_coo_tensor(torch.cat(idx, 1), torch.cat(val), (k * m, n),
                                   device=tensors[0].device, dtype=tensors[0].dtype).coalesce()


def combine_facet_operators(A_main, A_bc, b_bc, bc_edge_mask, n_edges, n_cells, n_comp, device):
    """Scatter interior-facet and boundary-facet operator rows into global facet order
    (sparse.py:176-228).  Returns (CSR operator, dense offset vector)."""
    main_g = torch.where(~bc_edge_mask)[0]
    bc_g = torch.where(bc_edge_mask)[0]
    mi, mv = A_main._indices(), A_main._values()
    bi, bv = A_bc._indices(), A_bc._values()
    rows = torch.cat([main_g[mi[0] // n_comp] * n_comp + mi[0] % n_comp, bc_g[bi[0] // n_comp] * n_comp + bi[0] % n_comp])
    A_all = torch.sparse_coo_tensor(torch.stack([rows, torch.cat([mi[1], bi[1]])]), torch.cat([mv, bv]),
                                    size=(n_edges * n_comp, n_cells * n_comp), device=device,
                                    dtype=A_main.dtype).coalesce()
    b_all = torch.zeros(n_edges * n_comp, device=device, dtype=b_bc.dtype)
    r = torch.arange(b_bc.numel(), device=device)
    b_all[bc_g[r // n_comp] * n_comp + r % n_comp] = b_bc
    return A_all.to_sparse_csr(), b_all
