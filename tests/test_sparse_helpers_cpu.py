"""The operator-assembly helpers of utils/sparse.py (reference utils/sparse.py:72-325) on
the CPU, each against a brute-force dense statement of what it must build:

* lift_sparse_matrix: A_new[i*n + c, j*n + c] = A[i, j], so A_new @ U.flatten() equals
  (A @ U).flatten() for any U [N, n] (sparse.py:231-278);
* interleave_sparse_rows: out[k*r + i] = tensors[i][r], with the reference's errors for
  an empty list, a non-COO input and mismatched shapes (sparse.py:281-325);
* combine_facet_operators: global facet e, component c takes row (j*n + c) of A_main
  when e is the j-th interior facet and row (j*n + c) of A_bc (plus offset b_bc[j*n + c])
  when e is the j-th boundary facet; interior rows get a zero offset (sparse.py:72-228).
"""
import numpy as np
import pytest
import torch

from fvm_solver_amd.utils.sparse import combine_facet_operators, interleave_sparse_rows, lift_sparse_matrix


def _rand_coo(m, n, density, rng, dtype=torch.float32):
    d = rng.standard_normal((m, n)) * (rng.random((m, n)) < density)
    return torch.from_numpy(d).to(dtype).to_sparse_coo(), d


@pytest.mark.parametrize("m,n,nc", [(1, 1, 1), (7, 5, 3), (40, 33, 4), (12, 60, 5)])
def test_lift_sparse_matrix_dense(m, n, nc):
    rng = np.random.default_rng(m * 100 + n)
    A, d = _rand_coo(m, n, 0.2, rng, torch.float64)
    L = lift_sparse_matrix(A, nc)
    assert L.layout == torch.sparse_coo and tuple(L.shape) == (m * nc, n * nc)
    ref = np.zeros((m * nc, n * nc))
    for i in range(m):
        for j in range(n):
            for c in range(nc):
                ref[i * nc + c, j * nc + c] = d[i, j]
    assert np.array_equal(L.to_dense().numpy(), ref)
    U = torch.from_numpy(rng.standard_normal((n, nc)))
    lhs = torch.sparse.mm(L, U.reshape(-1, 1)).reshape(m, nc)
    assert torch.allclose(lhs, torch.sparse.mm(A, U), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("k,m,n", [(1, 4, 4), (2, 9, 6), (3, 5, 11), (4, 1, 3)])
def test_interleave_sparse_rows_dense(k, m, n):
    rng = np.random.default_rng(k * 10 + m)
    mats = [_rand_coo(m, n, 0.4, rng) for _ in range(k)]
    out = interleave_sparse_rows([t for t, _ in mats])
    assert out.layout == torch.sparse_coo and tuple(out.shape) == (k * m, n) and out.dtype == torch.float32
    ref = np.zeros((k * m, n), np.float32)
    for r in range(m):
        for i, (_, d) in enumerate(mats):
            ref[k * r + i] = d[r]
    assert np.array_equal(out.to_dense().numpy(), ref)


def test_interleave_sparse_rows_errors():
    with pytest.raises(ValueError):
        interleave_sparse_rows([])
    a = torch.eye(3).to_sparse_coo()
    with pytest.raises(TypeError):
        interleave_sparse_rows([a, torch.eye(3)])
    with pytest.raises(ValueError):
        interleave_sparse_rows([a, torch.eye(4).to_sparse_coo()])


@pytest.mark.parametrize("n_edges,n_cells,nc,seed", [(6, 4, 1, 0), (20, 9, 4, 1), (33, 15, 3, 2), (5, 5, 2, 3)])
def test_combine_facet_operators_dense(n_edges, n_cells, nc, seed):
    rng = np.random.default_rng(seed)
    mask = rng.random(n_edges) < 0.35
    mask[0] = False
    mask[-1] = True
    nm, nb = int((~mask).sum()), int(mask.sum())
    Am, dm = _rand_coo(nm * nc, n_cells * nc, 0.3, rng)
    Ab, db = _rand_coo(nb * nc, n_cells * nc, 0.3, rng)
    bbc = torch.from_numpy(rng.standard_normal(nb * nc).astype(np.float32))
    A_all, b_all = combine_facet_operators(Am, Ab, bbc, torch.from_numpy(mask), n_edges, n_cells, nc,
                                           torch.device("cpu"))
    assert A_all.layout == torch.sparse_csr and tuple(A_all.shape) == (n_edges * nc, n_cells * nc)
    ref = np.zeros((n_edges * nc, n_cells * nc), np.float32)
    refb = np.zeros(n_edges * nc, np.float32)
    jm = jb = 0
    for e in range(n_edges):
        for c in range(nc):
            if mask[e]:
                ref[e * nc + c] = db[jb * nc + c]
                refb[e * nc + c] = bbc[jb * nc + c]
            else:
                ref[e * nc + c] = dm[jm * nc + c]
        if mask[e]:
            jb += 1
        else:
            jm += 1
    assert np.array_equal(A_all.to_dense().numpy(), ref)
    assert np.array_equal(b_all.numpy(), refb)
