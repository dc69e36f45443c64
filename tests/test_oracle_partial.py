"""The partial-spectrum oracle and the sparse containers on the CPU.

* the oracle's selection restates test/test_eigen.jl's expectations: LM = the num_bands
  eigenvalues nearest sigma (sort(e, by = abs)[1:nb] at sigma = 0), SR / LR the ends, always
  ascending -- checked against numpy's eigvalsh on random Hermitian matrices, with a planted
  mistake (selecting by e - sigma instead of |e - sigma|) that must differ;
* SparseHops: H(k) and the merged CSR against the dense Hops they came from;
* the sparse twisted-bilayer generator gives exactly the dense generator's blocks.
"""
import importlib

import numpy as np
import pytest


@pytest.fixture(scope="module")
def sparse(pkg):
    return importlib.import_module(pkg.__name__ + ".sparse")


def test_selection_matches_reference_semantics(pkg, oracle):
    from oracle import partial

    rng = np.random.default_rng(11)
    A = rng.standard_normal((60, 60)) - 0.5
    A = (A + A.T) / 2
    h = pkg.Hops({(0,): A})
    R, T = h.pack()
    ev = np.linalg.eigvalsh(A)
    ks = np.zeros((1, 1))
    for sigma in (0.0, 0.2, -1.3):
        b = partial.eigen_partial(R, T, ks, 10, sigma, "LM")[:, 0]
        ref = np.sort(ev[np.argsort(np.abs(ev - sigma))[:10]])
        assert np.max(np.abs(b - ref)) < 1e-12
        wrong = np.sort(ev[np.argsort(ev - sigma)[:10]])
        assert np.max(np.abs(wrong - ref)) > 1e-3
    assert np.max(np.abs(partial.eigen_partial(R, T, ks, 6, which="SR")[:, 0] - ev[:6])) < 1e-12
    assert np.max(np.abs(partial.eigen_partial(R, T, ks, 6, which="LR")[:, 0] - ev[-6:])) < 1e-12
    b, V = partial.eigen_partial(R, T, ks, 4, 0.3, "LM", vectors=True)
    for j in range(4):
        assert np.linalg.norm(A @ V[:, j, 0] - b[j, 0] * V[:, j, 0]) < 1e-10


def test_sparsehops_csr_and_bloch(pkg, sparse):
    h = pkg.models.random_tb(7, rmax=1, seed=3)
    sh = sparse.SparseHops.from_hops(h)
    R, rowptr, col, key, val = sh.csr()
    assert rowptr[0] == 0 and rowptr[-1] == col.size == sh.nnz
    # the library reads key r's components contiguously (column-major d x nR)
    flat = np.frombuffer(R.tobytes(order="A"), dtype=np.int32) if R.flags.f_contiguous else None
    assert flat is not None and [tuple(flat[2 * r:2 * r + 2]) for r in range(R.shape[1])] == list(sh.keys())
    k = np.array([0.13, -0.41])
    ph = np.exp(2j * np.pi * (k @ R))
    Hk = np.zeros((7, 7), complex)
    for i in range(7):
        for e in range(rowptr[i], rowptr[i + 1]):
            Hk[i, col[e]] += val[e] * ph[key[e]]
    assert np.max(np.abs(Hk - h(k))) < 1e-13
    assert np.max(np.abs(sh(k).toarray() - h(k))) < 1e-13
    assert sh.ishermitian(1e-14)


@pytest.mark.parametrize("m", [2, 3])
def test_sparse_tbg_equals_dense(pkg, m):
    hd = pkg.models.tbg_slater_koster(m)
    hs = pkg.models.tbg_slater_koster_sparse(m)
    assert set(hd.keys()) == set(hs.keys())
    for R in hd.keys():
        assert np.array_equal(hs[R].toarray(), hd[R])
