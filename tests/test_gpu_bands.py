"""GPU eigenvalues / eigenvectors vs the FP64 oracle (LAPACK zheevr) and closed forms.

Tolerance: |e_gpu - e_oracle| <= 1e-10 * max(|e|, ||H(k)||) per eigenvalue (the
eigenvalue bar, with the norm floor needed at e ~ 0); eigenvector residuals and
orthogonality at FP64 backward-stable level.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hnorm(h, ks):
    return np.array([np.linalg.norm(h(k), 2) for k in ks.T])


def _check_evals(b, ref, scale):
    err = np.abs(b - ref) / np.maximum(np.abs(ref), scale[None, :])
    assert np.max(err) < 1e-10, np.max(err)


def test_graphene_analytic(pkg, gpu):
    h = pkg.models.honeycomb_nn(-1.0)
    rng = np.random.default_rng(0)
    ks = rng.random((2, 200))
    b, _ = pkg.api.bandmatrix(h, ks)
    rs = [(1, 0), (-1, 0), (0, 1), (0, -1), (1, -1), (-1, 1)]
    up = np.sqrt(3 + sum(np.cos(2 * np.pi * (ks[0] * r[0] + ks[1] * r[1])) for r in rs))
    assert np.sum(np.abs(b[0] + up)) + np.sum(np.abs(b[1] - up)) < 1e-12 * 200


def test_square_triangular_analytic(pkg, gpu):
    rng = np.random.default_rng(1)
    ks = rng.random((2, 100))
    b, _ = pkg.api.bandmatrix(pkg.models.square_nn(-1.0), ks)
    assert np.max(np.abs(b[0] + 2 * (np.cos(2 * np.pi * ks[0]) + np.cos(2 * np.pi * ks[1])))) < 1e-12
    b, _ = pkg.api.bandmatrix(pkg.models.triangular_nn(-1.0), ks)
    tri = -2 * (np.cos(2 * np.pi * ks[0]) + np.cos(2 * np.pi * ks[1]) + np.cos(2 * np.pi * (ks[0] - ks[1])))
    assert np.max(np.abs(b[0] - tri)) < 1e-12


@pytest.fixture(params=[(0, -1), (1, -1), (2, -1), (3, -1), (4, -1), (0, 0)],
                ids=["tri-block", "tri-wave", "tri-reg", "tri-col", "tri-col2", "bt-loop"])
def tri_variant(request, gpu):
    """Tridiagonalization variant and back-transformation variant (-1: default, WY)."""
    gpu.set_variant(0, request.param[0])
    gpu.set_variant(4, request.param[1])
    yield request.param
    gpu.set_variant(0, -1)
    gpu.set_variant(4, -1)


@pytest.mark.parametrize("N,rmax,nk", [(1, 1, 9), (2, 1, 16), (3, 2, 16), (8, 1, 64), (17, 2, 25),
                                       (64, 2, 256), (65, 1, 16), (96, 1, 8), (128, 1, 8)])
def test_random_model_vs_oracle(pkg, oracle, gpu, tri_variant, N, rmax, nk):
    h = pkg.models.random_tb(N, rmax=rmax, seed=100 + N)
    ks = pkg.regulargrid(nk=nk)
    R, T = h.pack()
    ref = oracle.bands(R, T, ks)
    b, _ = pkg.api.bandmatrix(h, ks)
    _check_evals(b, ref, _hnorm(h, ks))


@pytest.mark.parametrize("N", [4, 30, 64, 100])
def test_eigenvectors(pkg, gpu, tri_variant, N):
    h = pkg.models.random_tb(N, rmax=1, seed=N)
    ks = pkg.regulargrid(nk=9)
    b, V = pkg.api.eigen(h, ks)
    for j in range(ks.shape[1]):
        hk = h(ks[:, j])
        nrm = np.linalg.norm(hk, 2)
        v = V[:, :, j]
        assert np.max(np.abs(hk @ v - v * b[:, j])) < 1e-12 * N * nrm
        assert np.max(np.abs(v.conj().T @ v - np.eye(N))) < 1e-12 * N


def test_degenerate_spinful_bands(pkg, oracle, gpu):
    # kron(h, sigma_0): every level exactly two-fold degenerate
    h = pkg.models.addspin(pkg.models.random_tb(6, rmax=1, seed=5), "spinhalf")
    ks = pkg.regulargrid(nk=16)
    b, V = pkg.api.eigen(h, ks)
    R, T = h.pack()
    _check_evals(b, oracle.bands(R, T, ks), _hnorm(h, ks))
    for j in range(ks.shape[1]):
        v = V[:, :, j]
        assert np.max(np.abs(v.conj().T @ v - np.eye(12))) < 1e-12
        assert np.max(np.abs(h(ks[:, j]) @ v - v * b[:, j])) < 1e-11


def test_upper_triangle_semantics(pkg, oracle, gpu):
    # a non-Hermitian onsite block: only its upper triangle counts (Hermitian(M))
    h = pkg.models.random_tb(5, rmax=1, seed=9)
    h[(0, 0)] = h[(0, 0)] + np.tril(np.ones((5, 5)), -1) * (0.3 + 0.2j)
    ks = pkg.regulargrid(nk=9)
    R, T = h.pack()
    ref = oracle.bands(R, T, ks)
    b, _ = pkg.api.bandmatrix(h, ks)
    assert np.max(np.abs(b - ref)) < 1e-11


def test_fouriersum(pkg, oracle, gpu):
    h = pkg.models.random_tb(7, rmax=2, seed=3)
    rng = np.random.default_rng(2)
    ks = rng.random((2, 5))
    Hk = pkg.api.fouriersum(h, ks)
    R, T = h.pack()
    for j in range(5):
        assert np.max(np.abs(Hk[:, :, j] - oracle.fouriersum(R, T, ks[:, j]))) < 1e-13


def test_single_kpoint_and_empty(pkg, gpu):
    h = pkg.models.random_tb(4, rmax=1, seed=1)
    b, _ = pkg.api.bandmatrix(h, np.zeros((2, 1)))
    assert np.allclose(b[:, 0], np.linalg.eigvalsh(h(np.zeros(2))), atol=1e-12)
    b, _ = pkg.api.bandmatrix(h, np.zeros((2, 0)))
    assert b.shape == (4, 0)


def test_nonfinite_input_raises(pkg, gpu):
    h = pkg.models.random_tb(4, rmax=1, seed=1)
    h[(0, 0)] = h[(0, 0)] * np.nan
    with pytest.raises((FloatingPointError, RuntimeError)):
        pkg.api.bandmatrix(h, pkg.regulargrid(nk=4))


def test_dimension_mismatch_raises(pkg, gpu):
    h = pkg.models.random_tb(4, rmax=1, seed=1)
    with pytest.raises(ValueError):
        pkg.api.bandmatrix(h, np.zeros((3, 2)))


def test_band_expectation_values(pkg, oracle, gpu):
    # projector expectation values real(psi^H O psi) (Spectrum/bands.jl:52-58, :107-112):
    # a k-independent matrix and a k-dependent Hops operator, against the oracle's eigenvectors
    N = 12
    h = pkg.models.random_tb(N, rmax=1, seed=31)
    ks = pkg.regulargrid(nk=25)
    rng = np.random.default_rng(8)
    M = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    O = pkg.models.random_tb(N, rmax=1, seed=32)
    b, obs = pkg.api.bandmatrix(h, ks, [M, O])
    R, T = h.pack()
    rb, V = oracle.bands(R, T, ks, vectors=True)
    _check_evals(b, rb, _hnorm(h, ks))
    assert obs.shape == (N, ks.shape[1], 2)
    for j in range(ks.shape[1]):
        v = V[:, :, j]
        e0 = np.real(np.einsum("in,ij,jn->n", v.conj(), M, v))
        e1 = np.real(np.einsum("in,ij,jn->n", v.conj(), O(ks[:, j]), v))
        assert np.max(np.abs(obs[:, j, 0] - e0)) < 1e-11
        assert np.max(np.abs(obs[:, j, 1] - e1)) < 1e-11
    bd = pkg.api.getbands(h, ks, M)
    assert bd.obs.shape == (N, ks.shape[1], 1)


def test_band_expectation_spin_zeeman(pkg, gpu):
    # s_z of spinful graphene in a z Zeeman field: every band has <s_z> = +-1
    h = pkg.models.addspin(pkg.models.honeycomb_nn(-1.0), "spinhalf")
    sz = np.kron(np.eye(2), np.diag([1.0, -1.0]))
    h[(0, 0)] = h[(0, 0)] + 0.3 * sz
    ks = np.random.default_rng(4).random((2, 50))
    b, obs = pkg.api.bandmatrix(h, ks, sz)
    assert np.max(np.abs(np.abs(obs[:, :, 0]) - 1.0)) < 1e-12
    assert np.max(np.abs(np.sum(obs[:, :, 0], axis=0))) < 1e-12
