"""Per-k-point tridiagonal eigenpairs (variant kind 6): bisection with Newton steps,
twisted-factorisation eigenvectors, Gram-Schmidt inside eigenvalue clusters and the QL path
for k-points whose cluster vectors come out dependent -- vs the FP64 oracle (LAPACK zheevr)
and vs the QL + rotation replay path (kind 6 variant 0).

Spectra with designed clusters: pairs at gaps 3e-3, 1e-6 and 1e-10, a triple within 2e-7,
an exact double level, and the spin-degenerate model (every level exactly twice).
Bars: eigenvalues 1e-10 relative to max(|e|, ||H||); eigenvectors: residual
1e-12 N ||H||, ||V^H V - I|| <= 1e-12 N; rho 1e-9, E_band 1e-10 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hnorm(h, ks):
    return np.array([np.linalg.norm(h(k), 2) for k in ks.T])


def _clustered_model(pkg, N, seed):
    """Onsite-only model H = Q diag(lam) Q^H with clusters designed into lam."""
    rng = np.random.default_rng(seed)
    lam = np.sort(rng.uniform(-3.0, 3.0, N))
    lam[5] = lam[4] + 3e-3
    lam[10] = lam[9] + 1e-6
    lam[20] = lam[19] + 1e-10
    lam[30] = lam[29]
    lam[38] = lam[37] + 1e-7
    lam[39] = lam[37] + 2e-7
    Q, _ = np.linalg.qr(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    H0 = (Q * lam) @ Q.conj().T
    return pkg.Hops({(0, 0): 0.5 * (H0 + H0.conj().T)})


def _models(pkg):
    return {
        "clustered48": _clustered_model(pkg, 48, 3),
        "spin2x24": pkg.models.addspin(pkg.models.random_tb(24, rmax=1, seed=9), "spinhalf"),
        "random64": pkg.models.random_tb(64, rmax=2, seed=21),
    }


@pytest.fixture(params=[1, 0], ids=["bisect-twist", "ql-replay"])
def trieig(request, gpu):
    gpu.set_variant(6, request.param)
    yield request.param
    gpu.set_variant(6, -1)


@pytest.mark.parametrize("name", ["clustered48", "spin2x24", "random64"])
def test_tri_eigenpairs_vs_oracle(pkg, oracle, gpu, trieig, name):
    h = _models(pkg)[name]
    N = h[(0, 0)].shape[0]
    ks = pkg.regulargrid(nk=9)
    b, V = pkg.api.eigen(h, ks)
    R, T = h.pack()
    ref = oracle.bands(R, T, ks)
    hn = _hnorm(h, ks)
    assert np.max(np.abs(b - ref) / np.maximum(np.abs(ref), hn[None, :])) < 1e-10
    for j in range(ks.shape[1]):
        hk = h(ks[:, j])
        v = V[:, :, j]
        assert np.max(np.abs(hk @ v - v * b[:, j])) < 1e-12 * N * hn[j]
        assert np.max(np.abs(v.conj().T @ v - np.eye(N))) < 1e-12 * N


@pytest.mark.parametrize("name", ["clustered48", "spin2x24", "random64"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tri_rho_vs_oracle(pkg, oracle, gpu, trieig, name, precision):
    h = _models(pkg)[name]
    N = h[(0, 0)].shape[0]
    ks = pkg.regulargrid(nk=16)
    ex = gpu if precision == "fp64" else pkg.api.to_executor("gpu", precision="mixed")
    ex.set_variant(6, trieig)
    try:
        rho = pkg.Hops({key: np.zeros((N, N), complex) for key in h.keys()})
        eb = pkg.api.getdensitymatrix(rho, h, ks, 0.1, T=0.02, executor=ex)
    finally:
        ex.set_variant(6, -1)
    R, Tt = h.pack()
    ref, eb_ref = oracle.densitymatrix(R, Tt, ks, R, mu=0.1, Temp=0.02)
    err = max(np.max(np.abs(rho[tuple(R[:, r])] - ref[:, :, r])) for r in range(R.shape[1]))
    assert err < 1e-9, err
    assert abs(eb - eb_ref) <= 1e-10 * max(1.0, abs(eb_ref))


@pytest.mark.parametrize("N", [5, 12, 17, 31, 33, 40, 48, 49, 63])
def test_tri_small_register_tiles(pkg, oracle, gpu, N):
    # N <= 16 reaches the twisted kernel's 8- and 16-entry instantiations only with the
    # compact-WY back-transformation forced on; 33, 48, 49 and 63 start the register-tile
    # tridiagonalization on 3x3 / 4x4 tiles and hand the trailing matrix to the next smaller
    # tiles after 1, 16, 1 and 15 columns
    gpu.set_variant(4, 1)
    try:
        h = pkg.models.random_tb(N, rmax=1, seed=40 + N)
        ks = pkg.regulargrid(nk=16)
        b, V = pkg.api.eigen(h, ks)
    finally:
        gpu.set_variant(4, -1)
    R, T = h.pack()
    ref = oracle.bands(R, T, ks)
    hn = _hnorm(h, ks)
    assert np.max(np.abs(b - ref) / np.maximum(np.abs(ref), hn[None, :])) < 1e-10
    for j in range(ks.shape[1]):
        v = V[:, :, j]
        assert np.max(np.abs(h(ks[:, j]) @ v - v * b[:, j])) < 1e-12 * N * hn[j]
        assert np.max(np.abs(v.conj().T @ v - np.eye(N))) < 1e-12 * N


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("var", [3, 4])
@pytest.mark.parametrize("N", [1, 2, 3, 9, 16, 17, 31, 32, 33, 40, 47, 48, 63, 64])
def test_tri_column_per_lane_sizes(pkg, oracle, var, prec, N):
    # column-per-lane tridiagonalization at the sizes where its phases (row chunks of 8) and
    # its wave split (32 rows per wave) change: eigenpairs and the density matrix vs the
    # oracle, in FP64 and in FP32 with refinement
    ex = pkg.api.to_executor("gpu", precision=prec)
    ex.set_variant(0, var)
    ex.set_variant(4, 1)
    try:
        h = pkg.models.random_tb(N, rmax=1, seed=70 + N)
        ks = pkg.regulargrid(nk=16)
        b, V = pkg.api.eigen(h, ks, executor=ex)
        rho = pkg.Hops({Rk: np.zeros((N, N), complex) for Rk in h.keys()})
        eb = pkg.api.getdensitymatrix(rho, h, ks, 0.05, T=0.02, executor=ex)
    finally:
        ex.set_variant(0, -1)
        ex.set_variant(4, -1)
    R, T = h.pack()
    ref = oracle.bands(R, T, ks)
    hn = _hnorm(h, ks)
    assert np.max(np.abs(b - ref) / np.maximum(np.abs(ref), hn[None, :])) < 1e-10
    for j in range(ks.shape[1]):
        v = V[:, :, j]
        assert np.max(np.abs(h(ks[:, j]) @ v - v * b[:, j])) < 1e-11 * N * hn[j]
        assert np.max(np.abs(v.conj().T @ v - np.eye(N))) < 1e-11 * N
    rref, eb_ref = oracle.densitymatrix(R, T, ks, R, mu=0.05, Temp=0.02)
    err = max(np.max(np.abs(rho[tuple(R[:, r])] - rref[:, :, r])) for r in range(R.shape[1]))
    assert err < 1e-9, err
    assert abs(eb - eb_ref) <= 1e-10 * max(1.0, abs(eb_ref))


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("var", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("tiny", [1e-20, 1e-160])
def test_tiny_householder_column(pkg, oracle, var, prec, tiny):
    # a column whose entries square below the normal range (1e-20 in FP32, 1e-160 in FP64):
    # the Householder scalars come from the column rescaled by its largest entry instead of
    # overflowing 1 / |alpha - beta|^2 (the converged honeycomb Hubbard mean field has such
    # columns at its spin-degenerate k-points)
    t = tiny
    H0 = np.array([[1.0, 0.3, 0, 0], [0.3, 2.0, 1j * t, t], [0, -1j * t, 3.0, 0.2], [0, t, 0.2, 4.0]])
    h = pkg.Hops({(0, 0): H0, (1, 0): 0.1 * np.eye(4), (-1, 0): 0.1 * np.eye(4)})
    ks = pkg.regulargrid(nk=16)
    ex = pkg.api.to_executor("gpu", precision=prec)
    ex.set_variant(0, var)
    try:
        b, _ = pkg.api.bandmatrix(h, ks, executor=ex)
        rho = pkg.Hops({Rk: np.zeros((4, 4), complex) for Rk in h.keys()})
        eb = pkg.api.getdensitymatrix(rho, h, ks, 2.5, T=0.02, executor=ex)
    finally:
        ex.set_variant(0, -1)
    R, T = h.pack()
    ref = oracle.bands(R, T, ks)
    assert np.all(np.isfinite(b))
    assert np.max(np.abs(b - ref)) < 1e-10 * 5
    rref, eb_ref = oracle.densitymatrix(R, T, ks, R, mu=2.5, Temp=0.02)
    err = max(np.max(np.abs(rho[tuple(R[:, r])] - rref[:, :, r])) for r in range(R.shape[1]))
    assert err < 1e-9, err
    assert abs(eb - eb_ref) <= 1e-10 * max(1.0, abs(eb_ref))


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_tiny_householder_column_large_n(pkg, oracle, prec):
    # the same on the large-N path: two random 65-orbital blocks coupled by entries of 1e-160,
    # so the reduction meets a column of tiny entries after the first block
    rng = np.random.default_rng(5)
    n1 = 65
    B = [rng.normal(size=(n1, n1)) + 1j * rng.normal(size=(n1, n1)) for _ in range(2)]
    B = [0.5 * (b + b.conj().T) for b in B]
    C = 1e-160 * (rng.normal(size=(n1, n1)) + 1j * rng.normal(size=(n1, n1)))
    H0 = np.block([[B[0], C], [C.conj().T, B[1]]])
    h = pkg.Hops({(0, 0): H0, (1, 0): 0.1 * np.eye(2 * n1), (-1, 0): 0.1 * np.eye(2 * n1)})
    ks = pkg.regulargrid(nk=4)
    ex = pkg.api.to_executor("gpu", precision=prec)
    b, _ = pkg.api.bandmatrix(h, ks, executor=ex)
    R, T = h.pack()
    ref = oracle.bands(R, T, ks)
    assert np.all(np.isfinite(b))
    assert np.max(np.abs(b - ref)) < 1e-10 * np.max(np.abs(ref))


def test_tri_variants_agree(pkg, gpu):
    # the two tridiagonal eigensolvers give the same eigenvalues to rounding
    h = pkg.models.random_tb(50, rmax=2, seed=77)
    ks = pkg.regulargrid(nk=64)
    out = []
    for v in (1, 0):
        gpu.set_variant(6, v)
        try:
            out.append(pkg.api.eigen(h, ks)[0])
        finally:
            gpu.set_variant(6, -1)
    assert np.max(np.abs(out[0] - out[1])) < 1e-12 * np.max(_hnorm(h, ks))
