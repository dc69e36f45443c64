"""Partial spectrum on the GPU (format = :sparse, num_bands): Chebyshev-filtered subspace
iteration against the oracle's selection from the full LAPACK spectrum and against the
reference's own eigen_sparse tests (test/test_eigen.jl).

The reference's random matrices come from Julia's Xoshiro streams, which numpy cannot
reproduce: the fixtures here are matrices of the same shape and distribution from numpy
(rand(N, N) - 1/2 symmetrised, N = 60), with the reference's bars: eigenvalues to 1e-8, residuals
||H u - l u|| below 1e-9 (1e-8 for the ascending-order test), orthonormality ||U^H U - I|| below
1e-10, ascending order, identical results on a repeated call.  On top of them: eigenvalues to
1e-10 ||H(k)|| against the oracle for sparse random models and the twisted bilayer (m = 11,
1588 orbitals, the sparse generator), batches of k-points, all three selections.
"""
import importlib

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sparse(pkg):
    return importlib.import_module(pkg.__name__ + ".sparse")


@pytest.fixture(scope="module")
def partial():
    from oracle import partial as p

    return p


def _sym60(seed):
    rng = np.random.default_rng(seed)
    A = rng.random((60, 60)) - 0.5
    return (A + A.T) / 2


def test_reference_small_dense_fixture(sparse, gpu):
    # test_eigen.jl:11-38: LM at sigma = 0, SR, LR (num_bands 10), eigmin / eigmax
    A = _sym60(0x1234)
    ev = np.linalg.eigvalsh(A)
    M = sp.csr_matrix(A)
    e_lm, U = sparse.eigen_sparse_matrix(M, 10, sigma=0.0)
    assert np.max(np.abs(e_lm - np.sort(ev[np.argsort(np.abs(ev))[:10]]))) < 1e-8
    e_sr, _ = sparse.eigen_sparse_matrix(M, 10, which="SR")
    assert np.max(np.abs(e_sr - ev[:10])) < 1e-8
    e_lr, _ = sparse.eigen_sparse_matrix(M, 10, which="LR")
    assert np.max(np.abs(e_lr - ev[-10:])) < 1e-8
    assert abs(sparse.eigmin_sparse(M) - ev[0]) < 1e-8
    assert abs(sparse.eigmax_sparse(M) - ev[-1]) < 1e-8


def test_reference_residuals_orthonormality_tbg5(pkg, sparse, gpu):
    # test_eigen.jl:40-62: graphene on honeycomb_twisted(5) at the second point of a 4-point
    # k-path, 10 bands nearest the default sigma
    lat = importlib.import_module(pkg.__name__ + ".lattice")
    L = lat.honeycomb_twisted(5)
    h = lat.graphene(L)
    k = np.array([1 / 3, -1 / 3]) / 3.0
    H = sp.csr_matrix(h(k))
    e, U = sparse.eigen_sparse_matrix(H, 10)
    for b in range(10):
        assert np.linalg.norm(H @ U[:, b] - e[b] * U[:, b]) < 1e-9
    assert np.linalg.norm(U.conj().T @ U - np.eye(10), 2) < 1e-10
    ev = np.linalg.eigvalsh(H.toarray())
    assert np.max(np.abs(e - np.sort(ev[np.argsort(np.abs(ev - 1e-8))[:10]]))) < 1e-10 * np.max(np.abs(ev))


def test_reference_ascending_and_deterministic(sparse, gpu):
    # test_eigen.jl:68-86: A = Hermitian(A + A'), num_bands 6, LM at sigma = 0.2, SR, LR
    rng = np.random.default_rng(0x5eed)
    A = rng.random((60, 60)) - 0.5
    A = A + A.T
    M = sp.csr_matrix(A)
    for which, kw in (("LM", {"sigma": 0.2}), ("SR", {}), ("LR", {})):
        e, U = sparse.eigen_sparse_matrix(M, 6, which=which, **kw)
        assert np.all(np.diff(e) >= 0)
        for j in range(6):
            assert np.linalg.norm(A @ U[:, j] - e[j] * U[:, j]) < 1e-8
    e1, _ = sparse.eigen_sparse_matrix(M, 4)
    e2, _ = sparse.eigen_sparse_matrix(M, 4)
    assert np.array_equal(e1, e2)


def _sparse_random(pkg, sparse, N, seed, density=0.01, nkeys=2):
    # a Hermitian sparse model with the keys (0, 0), (+-1, 0) [, (0, +-1)]
    rng = np.random.default_rng(seed)

    def rnd():
        M = sp.random(N, N, density=density, random_state=rng, format="csr")
        return M + 1j * sp.random(N, N, density=density, random_state=rng, format="csr")

    h0 = rnd()
    blocks = {(0, 0): (h0 + h0.conj().T) / 2}
    for key in [(1, 0), (0, 1)][:nkeys]:
        t = rnd()
        blocks[key] = t
        blocks[tuple(-x for x in key)] = t.conj().T
    return sparse.SparseHops(blocks)


def _dense_pack(sh):
    R = np.array(list(sh.keys()), dtype=np.int32).T
    T = np.stack([sh[tuple(r)].toarray() for r in R.T], axis=2)
    return np.asfortranarray(R), np.asfortranarray(T)


@pytest.mark.parametrize("N,nb,which,sigma", [
    (300, 8, "LM", 0.0),
    (300, 8, "LM", 0.35),
    (300, 5, "SR", 0.0),
    (300, 5, "LR", 0.0),
    (800, 20, "LM", -0.1),
    (40, 12, "LM", 0.1),    # subspace = the whole space
])
def test_sparse_random_vs_oracle(pkg, sparse, partial, N, nb, which, sigma):
    sh = _sparse_random(pkg, sparse, N, seed=N + nb)
    assert sh.ishermitian(1e-15)
    ks = pkg.regulargrid(nk=9)
    b, V, it = sparse.eigen_sparse(sh, ks, nb, sigma=sigma, which=which, return_iters=True)
    assert np.all(it > 0)
    R, T = _dense_pack(sh)
    ref = partial.eigen_partial(R, T, ks, nb, sigma, which)
    scale = max(1.0, np.max(np.abs(ref)))
    assert np.max(np.abs(b - ref)) < 1e-10 * scale
    for j, k in enumerate(ks.T):
        Hk = sh(k)
        U = V[:, :, j]
        assert np.max(np.linalg.norm(Hk @ U - U * b[:, j], axis=0)) < 1e-9 * scale
        assert np.linalg.norm(U.conj().T @ U - np.eye(nb), 2) < 1e-10


def test_tbg_m11_sparse_vs_oracle(pkg, sparse, partial):
    sh = pkg.models.tbg_slater_koster_sparse(11)
    ks = np.array([[0.0, 0.0], [1 / 3, -1 / 3], [0.17, 0.41], [0.5, 0.0]]).T
    b, V = sparse.eigen_sparse(sh, ks, 10, sigma=0.0)
    R, T = _dense_pack(sh)
    ref = partial.eigen_partial(R, T, ks, 10, 0.0, "LM")
    hn = max(np.linalg.norm(sh(k).toarray(), 2) for k in ks.T)
    assert np.max(np.abs(b - ref)) < 1e-10 * hn
    for j, k in enumerate(ks.T):
        Hk = sh(k)
        assert np.max(np.linalg.norm(Hk @ V[:, :, j] - V[:, :, j] * b[:, j], axis=0)) < 1e-9
    # the dense Hops gives the same model, also through bandmatrix(format="sparse")
    b2 = sparse.bandmatrix_sparse(pkg.models.tbg_slater_koster(11), ks[:, :1], 10, sigma=0.0)
    assert np.max(np.abs(b2[:, 0] - b[:, 0])) < 1e-10 * hn
    b3, obs = pkg.api.bandmatrix(sh, ks[:, 1:2], format="sparse", num_bands=10, sigma=0.0)
    assert obs.shape == (10, 1, 0) and np.max(np.abs(b3[:, 0] - b[:, 1])) < 1e-10 * hn


def test_argument_errors(pkg, sparse, gpu):
    sh = _sparse_random(pkg, sparse, 50, seed=1)
    ks = np.zeros((2, 1))
    with pytest.raises(Exception):
        sparse.eigen_sparse(sh, ks, 0)
    with pytest.raises(Exception):
        sparse.eigen_sparse(sh, ks, 60)
    with pytest.raises(ValueError):
        sparse.eigen_sparse(sh, ks, 4, which="XX")
    with pytest.raises(Exception):
        sparse.eigen_sparse(_sparse_random(pkg, sparse, 500, seed=2), ks, 57)


def test_degenerate_window_and_edges(pkg, sparse, partial, gpu):
    # two identical decoupled copies of a random model: every eigenvalue exactly doubly
    # degenerate, so a window of an even number of bands is well defined; compare the
    # eigenvalues and the spectral projector of the window (gauge free) with the oracle
    rng = np.random.default_rng(9)
    n = 40
    t0 = sp.random(n, n, density=0.1, random_state=rng) + 1j * sp.random(n, n, density=0.1, random_state=rng)
    t0 = (t0 + t0.conj().T) / 2
    t1 = sp.random(n, n, density=0.05, random_state=rng) * (1 + 0.5j)
    blk = lambda a: sp.block_diag([a, a], format="csr")
    sh = sparse.SparseHops({(0,): blk(t0), (1,): blk(t1), (-1,): blk(t1).conj().T})
    ks = np.array([[0.0, 0.23, 0.5]])
    R, T = _dense_pack(sh)
    for sigma in (0.05, 100.0, -100.0):
        b, V = sparse.eigen_sparse(sh, ks, 6, sigma=sigma)
        ref, Vr = partial.eigen_partial(R, T, ks, 6, sigma, "LM", vectors=True)
        assert np.max(np.abs(b - ref)) < 1e-10 * max(1.0, np.max(np.abs(ref)))
        for j in range(ks.shape[1]):
            P, Pr = V[:, :, j] @ V[:, :, j].conj().T, Vr[:, :, j] @ Vr[:, :, j].conj().T
            assert np.max(np.abs(P - Pr)) < 1e-9
    # sigma far above the spectrum selects the largest eigenvalues, far below the smallest
    bl, _ = sparse.eigen_sparse(sh, ks, 4, sigma=100.0)
    br, _ = sparse.eigen_sparse(sh, ks, 4, which="LR")
    assert np.max(np.abs(bl - br)) < 1e-10
    # the whole spectrum (num_bands = N) and a 1 x 1 model
    small = _sparse_random(pkg, sparse, 20, seed=5, density=0.2)
    b, _ = sparse.eigen_sparse(small, np.array([[0.1], [0.3]]), 20)
    R2, T2 = _dense_pack(small)
    ref = partial.eigen_partial(R2, T2, np.array([[0.1], [0.3]]), 20, 1e-8, "LM")
    assert np.max(np.abs(b - ref)) < 1e-10 * max(1.0, np.max(np.abs(ref)))
    e, U = sparse.eigen_sparse_matrix(sp.csr_matrix([[0.7]]), 1)
    assert abs(e[0] - 0.7) < 1e-15 and abs(abs(U[0, 0]) - 1.0) < 1e-15
