"""Large-N eigensolver (blocked Householder tridiagonalization, Sturm bisection, inverse
iteration, compact-WY back-transformation) vs the FP64 oracle (LAPACK zheevr).

Bars: eigenvalues |e - e64| <= 1e-10 max(|e64|, ||H(k)||_2); rho to 1e-9 absolute;
E_band to 1e-10 relative; eigenvectors: residual ||H v - e v|| <= 1e-10 ||H||,
||V^H V - I|| <= 1e-10, and equal to the oracle's up to a phase where the eigenvalue is
isolated.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def blocked(pkg, gpu):
    """FP64 executor with the large-N path forced for every N."""
    ex = pkg.api.to_executor("gpu")
    ex.set_variant(5, 1)
    yield ex
    ex.set_variant(5, -1)


def _hnorm(h, ks):
    return np.array([np.linalg.norm(h(k), 2) for k in ks.T])


def _degenerate_model(pkg, n, seed):
    return pkg.models.addspin(pkg.models.random_tb(n, rmax=1, seed=seed), "spinhalf")


@pytest.mark.parametrize("N,nk", [(1, 4), (2, 9), (5, 16), (33, 9), (70, 9), (97, 4), (200, 4)])
def test_blocked_bands(pkg, oracle, blocked, N, nk):
    h = pkg.models.random_tb(N, rmax=1, seed=500 + N)
    ks = pkg.regulargrid(nk=nk)
    R, T = h.pack()
    ref = oracle.bands(R, T, ks)
    b, _ = pkg.api.bandmatrix(h, ks, executor=blocked)
    err = np.abs(b - ref) / np.maximum(np.abs(ref), _hnorm(h, ks)[None, :])
    assert np.max(err) < 1e-10, np.max(err)


@pytest.mark.parametrize("N,nk,mu,T", [(5, 16, 0.1, 0.05), (70, 9, 0.0, 0.02), (130, 4, 0.2, 0.0)])
def test_blocked_rho(pkg, oracle, blocked, N, nk, mu, T):
    h = pkg.models.random_tb(N, rmax=1, seed=600 + N)
    ks = pkg.regulargrid(nk=nk)
    rho = pkg.Hops({R: np.zeros((N, N), complex) for R in h.keys()})
    eb = pkg.api.getdensitymatrix(rho, h, ks, mu, T=T, executor=blocked)
    R, Tt = h.pack()
    ref, eb_ref = oracle.densitymatrix(R, Tt, ks, R, mu=mu, Temp=T)
    err = max(np.max(np.abs(rho[tuple(R[:, r])] - ref[:, :, r])) for r in range(R.shape[1]))
    assert err < 1e-9, err
    assert abs(eb - eb_ref) <= 1e-10 * max(1.0, abs(eb_ref))


def test_blocked_degenerate_spectrum(pkg, oracle, blocked):
    # exactly doubly degenerate levels (spin copies): clusters orthogonalised together
    h = _degenerate_model(pkg, 40, 9)
    ks = pkg.regulargrid(nk=4)
    R, Tt = h.pack()
    ref = oracle.bands(R, Tt, ks)
    b, _ = pkg.api.bandmatrix(h, ks, executor=blocked)
    assert np.max(np.abs(b - ref) / np.maximum(np.abs(ref), _hnorm(h, ks)[None, :])) < 1e-10
    rho = pkg.Hops({key: np.zeros((80, 80), complex) for key in h.keys()})
    eb = pkg.api.getdensitymatrix(rho, h, ks, 0.05, T=0.0, executor=blocked)
    refr, eb_ref = oracle.densitymatrix(R, Tt, ks, R, mu=0.05, Temp=0.0)
    err = max(np.max(np.abs(rho[tuple(R[:, r])] - refr[:, :, r])) for r in range(R.shape[1]))
    assert err < 1e-9, err
    assert abs(eb - eb_ref) <= 1e-10 * max(1.0, abs(eb_ref))


@pytest.mark.parametrize("N", [70, 150])
def test_blocked_eigenvectors(pkg, oracle, blocked, N):
    h = pkg.models.random_tb(N, rmax=1, seed=700 + N)
    ks = pkg.regulargrid(nk=4)
    b, V = pkg.api.eigen(h, ks, executor=blocked)
    R, Tt = h.pack()
    eb, Vr = oracle.bands(R, Tt, ks, vectors=True)
    for j in range(ks.shape[1]):
        H = h(ks[:, j])
        Hn = np.linalg.norm(H, 2)
        Vj = V[:, :, j]
        assert np.max(np.abs(H @ Vj - Vj * b[:, j])) < 1e-10 * Hn
        assert np.max(np.abs(Vj.conj().T @ Vj - np.eye(N))) < 1e-10
        # isolated eigenvalues: the same vector up to a phase
        gaps = np.diff(eb[:, j])
        iso = np.ones(N, bool)
        iso[:-1] &= gaps > 1e-6 * Hn
        iso[1:] &= gaps > 1e-6 * Hn
        ov = np.abs(np.sum(Vr[:, iso, j].conj() * Vj[:, iso], axis=0))
        assert np.max(np.abs(ov - 1.0)) < 1e-9


def test_blocked_densitymatrix_occ(pkg, oracle, blocked):
    N = 60
    h = pkg.models.random_tb(N, rmax=1, seed=88)
    ks = pkg.regulargrid(nk=9)
    R, Tt = h.pack()
    b, V = oracle.bands(R, Tt, ks, vectors=True)
    occ = np.random.default_rng(5).standard_normal(b.shape)
    rhos = pkg.Hops({key: np.zeros((N, N), complex) for key in [(0, 0), (1, 0), (1, -1)]})
    pkg.api.densitymatrix_occ(rhos, h, ks, occ, executor=blocked)
    for key in rhos.keys():
        ref = sum(np.exp(-2j * np.pi * np.dot(ks[:, j], key)) * (V[:, :, j] * occ[:, j]) @ V[:, :, j].conj().T
                  for j in range(ks.shape[1])) / ks.shape[1]
        assert np.max(np.abs(rhos[key] - ref)) < 1e-9


def test_default_path_above_512(pkg, oracle, gpu):
    # N > 512 takes the large-N path by default, in the precision the executor asks for
    N = 600
    h = pkg.models.random_tb(N, rmax=1, seed=3)
    ks = pkg.regulargrid(nk=4)
    R, Tt = h.pack()
    ref, eb_ref = oracle.densitymatrix(R, Tt, ks, R, mu=0.0, Temp=0.02)
    refb = oracle.bands(R, Tt, ks)
    for ex in (gpu, pkg.api.to_executor("gpu", precision="mixed")):
        b, _ = pkg.api.bandmatrix(h, ks, executor=ex)
        assert ex.last_precision() == ("fp64" if ex is gpu else "mixed")
        assert np.max(np.abs(b - refb) / np.maximum(np.abs(refb), _hnorm(h, ks)[None, :])) < 1e-10
        rho = pkg.Hops({key: np.zeros((N, N), complex) for key in h.keys()})
        eb = pkg.api.getdensitymatrix(rho, h, ks, 0.0, T=0.02, executor=ex)
        err = max(np.max(np.abs(rho[tuple(R[:, r])] - ref[:, :, r])) for r in range(R.shape[1]))
        assert err < 1e-9, err
        assert abs(eb - eb_ref) <= 1e-10 * max(1.0, abs(eb_ref))


def test_tbg_m11_slater_koster(pkg, oracle, gpu):
    # the 1588-orbital twisted bilayer (twist index 11) at Gamma and a generic k-point
    h = pkg.models.tbg_slater_koster(11)
    N = h[(0, 0)].shape[0]
    assert N == 1588
    ks = np.array([[0.0, 0.0], [0.137, 0.291]]).T
    R, Tt = h.pack()
    refb = oracle.bands(R, Tt, ks)
    b, _ = pkg.api.bandmatrix(h, ks, executor=gpu)
    hn = _hnorm(h, ks)
    assert np.max(np.abs(b - refb) / np.maximum(np.abs(refb), hn[None, :])) < 1e-10
    keys = [(0, 0), (1, 0), (0, 1)]
    Rr = np.array(keys).T
    ref, eb_ref = oracle.densitymatrix(R, Tt, ks, Rr, mu=0.0, Temp=0.01)
    rho = pkg.Hops({key: np.zeros((N, N), complex) for key in keys})
    eb = pkg.api.getdensitymatrix(rho, h, ks, 0.0, T=0.01, executor=gpu)
    err = max(np.max(np.abs(rho[key] - ref[:, :, r])) for r, key in enumerate(keys))
    assert err < 1e-9, err
    assert abs(eb - eb_ref) <= 1e-10 * max(1.0, abs(eb_ref))


def test_blocked_expectation_values(pkg, oracle, blocked):
    # projector expectation values on the large-N eigenvectors
    N = 70
    h = pkg.models.random_tb(N, rmax=1, seed=71)
    ks = pkg.regulargrid(nk=4)
    M = np.diag(np.linspace(-1.0, 1.0, N)).astype(complex)
    b, obs = pkg.api.bandmatrix(h, ks, M, executor=blocked)
    R, T = h.pack()
    rb, V = oracle.bands(R, T, ks, vectors=True)
    assert np.max(np.abs(b - rb)) < 1e-10 * np.max(_hnorm(h, ks))
    for j in range(ks.shape[1]):
        v = V[:, :, j]
        e0 = np.real(np.einsum("in,ij,jn->n", v.conj(), M, v))
        assert np.max(np.abs(obs[:, j, 0] - e0)) < 1e-10


def test_autotune_chooses_eigensolver_layout(pkg, oracle, gpu):
    # between the per-k kernels and the large-N path (kind 5) for 64 < N <= 512, on a
    # private context so the tuned choice does not leak into other tests
    import torch

    ex = pkg.api.GPUExec(0, pkg.api._PRECISIONS["fp64"])
    N = 96
    h = pkg.models.random_tb(N, rmax=1, seed=13)
    ks = pkg.regulargrid(nk=64)
    m = pkg.api.GPUModel(h, ex)
    ks_d = torch.from_numpy(np.ascontiguousarray(ks.T)).cuda()
    chosen = pkg.api.autotune(m, ks_d.data_ptr(), ks.shape[1])
    assert chosen[5] in (0, 1)
    rho = pkg.Hops({key: np.zeros((N, N), complex) for key in h.keys()})
    eb = pkg.api.getdensitymatrix(rho, h, ks, 0.0, T=0.05, executor=ex)
    R, Tt = h.pack()
    ref, eb_ref = oracle.densitymatrix(R, Tt, ks, R, mu=0.0, Temp=0.05)
    err = max(np.max(np.abs(rho[tuple(R[:, r])] - ref[:, :, r])) for r in range(R.shape[1]))
    assert err < 1e-9, err
    assert abs(eb - eb_ref) <= 1e-10 * max(1.0, abs(eb_ref))


# ---- the large-N path in mixed precision: FP32 reduction, FP64 Ogita-Aishima refinement ----
@pytest.fixture(scope="module")
def mixed_blocked(pkg, gpu):
    """Mixed-precision executor with the large-N path forced for every N."""
    ex = pkg.api.to_executor("gpu", precision="mixed")
    ex.set_variant(5, 1)
    yield ex
    ex.set_variant(5, -1)


@pytest.mark.parametrize("N,nk", [(1, 4), (2, 9), (33, 9), (97, 4), (200, 4)])
def test_mixed_blocked_bands(pkg, oracle, mixed_blocked, N, nk):
    h = pkg.models.random_tb(N, rmax=1, seed=500 + N)
    ks = pkg.regulargrid(nk=nk)
    R, T = h.pack()
    ref = oracle.bands(R, T, ks)
    b, _ = pkg.api.bandmatrix(h, ks, executor=mixed_blocked)
    err = np.abs(b - ref) / np.maximum(np.abs(ref), _hnorm(h, ks)[None, :])
    assert np.max(err) < 1e-10, np.max(err)
    assert mixed_blocked.last_precision() == "mixed"
    assert mixed_blocked.refinement_report()["sweeps"] >= 1


@pytest.mark.parametrize("N,nk,mu,T", [(5, 16, 0.1, 0.05), (70, 9, 0.0, 0.02), (130, 4, 0.2, 0.0)])
def test_mixed_blocked_rho(pkg, oracle, mixed_blocked, N, nk, mu, T):
    h = pkg.models.random_tb(N, rmax=1, seed=600 + N)
    ks = pkg.regulargrid(nk=nk)
    rho = pkg.Hops({R: np.zeros((N, N), complex) for R in h.keys()})
    eb = pkg.api.getdensitymatrix(rho, h, ks, mu, T=T, executor=mixed_blocked)
    R, Tt = h.pack()
    ref, eb_ref = oracle.densitymatrix(R, Tt, ks, R, mu=mu, Temp=T)
    err = max(np.max(np.abs(rho[tuple(R[:, r])] - ref[:, :, r])) for r in range(R.shape[1]))
    assert err < 1e-9, err
    assert abs(eb - eb_ref) <= 1e-10 * max(1.0, abs(eb_ref))
    rep = mixed_blocked.refinement_report()
    assert mixed_blocked.last_precision() == "mixed" and rep["max_orth_residual"] < 1e-3


def test_mixed_blocked_degenerate_and_occ(pkg, oracle, mixed_blocked):
    # exact spin degeneracy (clusters resolved by the Rayleigh-Ritz step), then given occupations
    h = _degenerate_model(pkg, 40, 9)
    ks = pkg.regulargrid(nk=4)
    R, Tt = h.pack()
    ref = oracle.bands(R, Tt, ks)
    b, _ = pkg.api.bandmatrix(h, ks, executor=mixed_blocked)
    assert np.max(np.abs(b - ref) / np.maximum(np.abs(ref), _hnorm(h, ks)[None, :])) < 1e-10
    rho = pkg.Hops({key: np.zeros((80, 80), complex) for key in h.keys()})
    eb = pkg.api.getdensitymatrix(rho, h, ks, 0.05, T=0.0, executor=mixed_blocked)
    refr, eb_ref = oracle.densitymatrix(R, Tt, ks, R, mu=0.05, Temp=0.0)
    err = max(np.max(np.abs(rho[tuple(R[:, r])] - refr[:, :, r])) for r in range(R.shape[1]))
    assert err < 1e-9, err
    assert abs(eb - eb_ref) <= 1e-10 * max(1.0, abs(eb_ref))
    N = 60
    h = pkg.models.random_tb(N, rmax=1, seed=88)
    ks = pkg.regulargrid(nk=9)
    R, Tt = h.pack()
    b, V = oracle.bands(R, Tt, ks, vectors=True)
    occ = np.random.default_rng(5).standard_normal(b.shape)
    rhos = pkg.Hops({key: np.zeros((N, N), complex) for key in [(0, 0), (1, 0), (1, -1)]})
    pkg.api.densitymatrix_occ(rhos, h, ks, occ, executor=mixed_blocked)
    for key in rhos.keys():
        ref = sum(np.exp(-2j * np.pi * np.dot(ks[:, j], key)) * (V[:, :, j] * occ[:, j]) @ V[:, :, j].conj().T
                  for j in range(ks.shape[1])) / ks.shape[1]
        assert np.max(np.abs(rhos[key] - ref)) < 1e-9


def test_tbg_m11_slater_koster_mixed(pkg, oracle, gpu):
    # the 1588-orbital twisted bilayer at Gamma and a generic k-point through the FP32 blocked
    # eigensolver and the FP64 refinement (the default layout above 64 orbitals)
    ex = pkg.api.to_executor("gpu", precision="mixed")
    h = pkg.models.tbg_slater_koster(11)
    N = h[(0, 0)].shape[0]
    ks = np.array([[0.0, 0.0], [0.137, 0.291]]).T
    R, Tt = h.pack()
    refb = oracle.bands(R, Tt, ks)
    b, _ = pkg.api.bandmatrix(h, ks, executor=ex)
    assert ex.last_precision() == "mixed"
    hn = _hnorm(h, ks)
    assert np.max(np.abs(b - refb) / np.maximum(np.abs(refb), hn[None, :])) < 1e-10
    keys = [(0, 0), (1, 0), (0, 1)]
    Rr = np.array(keys).T
    ref, eb_ref = oracle.densitymatrix(R, Tt, ks, Rr, mu=0.0, Temp=0.01)
    rho = pkg.Hops({key: np.zeros((N, N), complex) for key in keys})
    eb = pkg.api.getdensitymatrix(rho, h, ks, 0.0, T=0.01, executor=ex)
    err = max(np.max(np.abs(rho[key] - ref[:, :, r])) for r, key in enumerate(keys))
    assert err < 1e-9, err
    assert abs(eb - eb_ref) <= 1e-10 * max(1.0, abs(eb_ref))
    assert ex.refinement_report()["sweeps"] >= 1


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("N", [65, 66, 95, 96, 97, 127, 128])
def test_blocked_column_per_lane_reduction(pkg, oracle, prec, N):
    # 64 < N <= 128: the large-N path reduces H(k) with the column-per-lane kernel (two
    # column groups x four row groups); sizes at the column-group and row-group edges.
    # Eigenpairs and rho vs the oracle, and the eigenvalues vs the zlatrd panels (kind 0
    # variant 0 keeps them)
    ex = pkg.api.to_executor("gpu", precision=prec)
    ex.set_variant(5, 1)
    h = pkg.models.random_tb(N, rmax=1, seed=900 + N)
    ks = pkg.regulargrid(nk=9)
    R, T = h.pack()
    try:
        b, _ = pkg.api.bandmatrix(h, ks, executor=ex)
        rho = pkg.Hops({Rk: np.zeros((N, N), complex) for Rk in h.keys()})
        eb = pkg.api.getdensitymatrix(rho, h, ks, 0.05, T=0.02, executor=ex)
        ex.set_variant(0, 0)
        b0, _ = pkg.api.bandmatrix(h, ks, executor=ex)
        if prec == "fp64":
            ex.set_variant(0, -1)
            bv, V = pkg.api.eigen(h, ks, executor=ex)
    finally:
        ex.set_variant(0, -1)
        ex.set_variant(5, -1)
    ref = oracle.bands(R, T, ks)
    hn = _hnorm(h, ks)
    assert np.max(np.abs(b - ref) / np.maximum(np.abs(ref), hn[None, :])) < 1e-10
    assert np.max(np.abs(b - b0)) < 1e-10 * np.max(hn)
    rref, eb_ref = oracle.densitymatrix(R, T, ks, R, mu=0.05, Temp=0.02)
    err = max(np.max(np.abs(rho[tuple(R[:, r])] - rref[:, :, r])) for r in range(R.shape[1]))
    assert err < 1e-9, err
    assert abs(eb - eb_ref) <= 1e-10 * max(1.0, abs(eb_ref))
    if prec == "fp64":
        for j in range(ks.shape[1]):
            v = V[:, :, j]
            assert np.max(np.abs(h(ks[:, j]) @ v - v * bv[:, j])) < 1e-10 * hn[j]
            assert np.max(np.abs(v.conj().T @ v - np.eye(N))) < 1e-10


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_blocked_tiny_column_mid_n(pkg, oracle, prec):
    # a column of tiny entries in the column-per-lane reduction (two blocks coupled by 1e-20
    # in FP32 terms, 1e-160 in FP64): rescaled Householder scalars, no overflow
    rng = np.random.default_rng(8)
    n1 = 50
    B = [rng.normal(size=(n1, n1)) + 1j * rng.normal(size=(n1, n1)) for _ in range(2)]
    B = [0.5 * (b + b.conj().T) for b in B]
    t = 1e-20 if prec == "mixed" else 1e-160
    C = t * (rng.normal(size=(n1, n1)) + 1j * rng.normal(size=(n1, n1)))
    H0 = np.block([[B[0], C], [C.conj().T, B[1]]])
    h = pkg.Hops({(0, 0): H0, (1, 0): 0.1 * np.eye(2 * n1), (-1, 0): 0.1 * np.eye(2 * n1)})
    ks = pkg.regulargrid(nk=4)
    ex = pkg.api.to_executor("gpu", precision=prec)
    ex.set_variant(5, 1)
    try:
        b, _ = pkg.api.bandmatrix(h, ks, executor=ex)
    finally:
        ex.set_variant(5, -1)
    R, T = h.pack()
    ref = oracle.bands(R, T, ks)
    assert np.all(np.isfinite(b))
    assert np.max(np.abs(b - ref)) < 1e-10 * np.max(np.abs(ref))


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_blocked_split_reduction(pkg, oracle, prec):
    # a chunk of >= 64 k-points is reduced as two halves on two streams; with stage timing on it
    # runs on one stream: the same kernels on the same data, so the results agree bit for bit,
    # and both agree with the oracle (150 orbitals: the zlatrd panels; 81 k-points: halves of
    # 40 and 41)
    _lib = pkg.api._lib
    h = pkg.models.random_tb(150, rmax=1, seed=1150)
    ks = pkg.regulargrid(nk=81)
    ex = pkg.api.to_executor("gpu", precision=prec)
    out = []
    for timing in (0, 1):
        _lib.call("lqm_context_set_timing", ex.handle, timing)
        try:
            b, _ = pkg.api.bandmatrix(h, ks, executor=ex)
            rho = pkg.Hops({Rk: np.zeros((150, 150), complex) for Rk in h.keys()})
            eb = pkg.api.getdensitymatrix(rho, h, ks, 0.05, T=0.02, executor=ex)
        finally:
            _lib.call("lqm_context_set_timing", ex.handle, 0)
        out.append((b, rho, eb))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][2] == out[1][2]
    for Rk in h.keys():
        assert np.array_equal(out[0][1][Rk], out[1][1][Rk])
    R, T = h.pack()
    ref = oracle.bands(R, T, ks)
    assert np.max(np.abs(out[0][0] - ref) / np.maximum(np.abs(ref), _hnorm(h, ks)[None, :])) < 1e-10
    rref, eb_ref = oracle.densitymatrix(R, T, ks, R, mu=0.05, Temp=0.02)
    err = max(np.max(np.abs(out[0][1][tuple(R[:, r])] - rref[:, :, r])) for r in range(R.shape[1]))
    assert err < 1e-9, err
    assert abs(out[0][2] - eb_ref) <= 1e-10 * max(1.0, abs(eb_ref))
