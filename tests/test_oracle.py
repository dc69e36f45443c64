"""Pins the FP64 CPU oracle against closed forms, LAPACK via numpy, and the
invariants the reference's own tests assert.  CPU only."""
import numpy as np
import pytest


def _graphene_band(k):
    # test/analyticbands.jl:5-16
    rs = [(1, 0), (-1, 0), (0, 1), (0, -1), (1, -1), (-1, 1)]
    return np.sqrt(3.0 + sum(np.cos(2 * np.pi * (k[0] * r[0] + k[1] * r[1])) for r in rs))


def _rand_herm(rng, n):
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return (a + a.conj().T) / 2


@pytest.mark.parametrize("n", [1, 2, 5, 33])
def test_eigh_matches_numpy_lapack(oracle, n):
    rng = np.random.default_rng(n)
    A = _rand_herm(rng, n)
    w, V = oracle.eigh(A)
    wn = np.linalg.eigh(A)[0]
    assert np.max(np.abs(w - wn)) < 1e-12 * max(1.0, np.max(np.abs(wn)))
    assert np.max(np.abs(A @ V - V * w)) < 1e-12 * n
    assert np.max(np.abs(V.conj().T @ V - np.eye(n))) < 1e-13 * n


def test_eigh_reads_upper_triangle(oracle):
    # Hermitian(M) uses the upper triangle only (Eigen/Eigen.jl:25-31)
    rng = np.random.default_rng(3)
    A = _rand_herm(rng, 6)
    B = np.triu(A) + np.tril(rng.standard_normal((6, 6)), -1)
    assert np.allclose(oracle.eigh(B, vectors=False), np.linalg.eigvalsh(A, UPLO="U"), atol=1e-13)


def test_graphene_bands_analytic(pkg, oracle):
    # test_spectrum_graphene.jl:15-30: total MAE < 1e-12
    h = pkg.models.honeycomb_nn(-1.0)
    R, T = h.pack()
    rng = np.random.default_rng(0)
    ks = rng.random((2, 200))
    b = oracle.bands(R, T, ks)
    up = np.array([_graphene_band(k) for k in ks.T])
    assert np.sum(np.abs(b[0] + up)) + np.sum(np.abs(b[1] - up)) < 1e-12 * 200


def test_square_and_triangular_analytic(pkg, oracle):
    # test_spectrum_square.jl:10-20, test_spectrum_triangular.jl:10-33
    rng = np.random.default_rng(1)
    ks = rng.random((2, 300))
    R, T = pkg.models.square_nn(-1.0).pack()
    sq = -2 * (np.cos(2 * np.pi * ks[0]) + np.cos(2 * np.pi * ks[1]))
    assert np.max(np.abs(oracle.bands(R, T, ks)[0] - sq)) < 1e-12
    R, T = pkg.models.triangular_nn(-1.0).pack()
    tri = -2 * (np.cos(2 * np.pi * ks[0]) + np.cos(2 * np.pi * ks[1]) + np.cos(2 * np.pi * (ks[0] - ks[1])))
    assert np.max(np.abs(oracle.bands(R, T, ks)[0] - tri)) < 1e-12
    gam_k = oracle.bands(R, T, np.array([[0.0, 1 / 3], [0.0, 2 / 3]]))[0]
    assert abs(gam_k[0] + 6) < 1e-12 and abs(gam_k[1] - 3) < 1e-12


def test_graphene_dos_invariants(pkg, oracle):
    # test_spectrum_graphene.jl:42-63
    h = pkg.models.honeycomb_nn(-1.0)
    R, T = h.pack()
    om = np.linspace(-3.05, 3.05, 401)
    ks = pkg.regulargrid(nk=120 ** 2)
    dos = oracle.dos(R, T, ks, om, 0.05)
    assert np.max(np.abs(dos - dos[::-1])) / np.max(np.abs(dos)) < 1e-10
    assert abs(np.sum(dos) * (om[1] - om[0]) - 2 * np.pi) < 0.05 * 2 * np.pi
    i0, ip, im = np.argmin(np.abs(om)), np.argmin(np.abs(om - 1)), np.argmin(np.abs(om + 1))
    assert dos[ip] > 3 * dos[i0] and dos[im] > 3 * dos[i0]


def test_dos_is_lorentzian_sum_over_bands(pkg, oracle):
    h = pkg.models.random_tb(5, rmax=1, seed=4)
    R, T = h.pack()
    ks = pkg.regulargrid(nk=16)
    om = np.linspace(-4, 4, 33)
    b = oracle.bands(R, T, ks)
    g = 0.1
    ref = np.sum(g / ((om[:, None] - b.ravel()[None, :]) ** 2 + g * g), axis=1) / ks.shape[1]
    assert np.max(np.abs(oracle.dos(R, T, ks, om, g) - ref)) < 1e-12 * np.max(ref)


def _zeeman_graphene(pkg, B=0.4):
    h = pkg.models.addspin(pkg.models.honeycomb_nn(-1.0), "spinhalf")
    sy = np.array([[0, -1j], [1j, 0]])
    h[(0, 0)] = h[(0, 0)] + np.kron(np.eye(2), B * sy)
    return h, np.kron(np.eye(2), sy)


def test_densitymatrix_vs_direct_eigen(pkg, oracle):
    # test_operators.jl:91-127: expval(rho, O) = (1/Nk) sum_k tr[rho(k) O(k)], L pairs with -L
    h, Sy = _zeeman_graphene(pkg)
    nk = 9
    ks = np.array([[i / nk, j / nk] for i in range(nk) for j in range(nk)]).T
    mu = -0.5
    R, T = h.pack()
    rho, eb = oracle.densitymatrix(R, T, ks, R, mu=mu, Temp=0.0)
    sy_direct = eband_direct = 0.0
    for k in ks.T:
        hk = h(k)
        e, U = np.linalg.eigh(hk)
        rk = (U * (e < mu)) @ U.conj().T
        sy_direct += np.real(np.trace(rk @ Sy))
        eband_direct += np.real(np.trace(rk @ hk))
    sy_direct /= nk * nk
    eband_direct /= nk * nk
    keys = [tuple(R[:, r]) for r in range(R.shape[1])]
    i0 = keys.index((0, 0))
    assert abs(sy_direct) > 1e-3
    assert abs(np.real(np.trace(rho[:, :, i0] @ Sy)) - sy_direct) < 1e-10
    pair = sum(np.trace(rho[:, :, r] @ h[tuple(-x for x in keys[r])]) for r in range(len(keys)))
    assert abs(np.real(pair) - eband_direct) < 1e-10
    assert abs(eb - eband_direct) < 1e-10
    # planted mistake: pairing L with +L (the old bug the reference test guards) fails
    wrong = sum(np.trace(rho[:, :, r] @ h[keys[r]]) for r in range(len(keys)))
    assert abs(np.real(wrong) - eband_direct) > 1e-3


def test_densitymatrix_hermitian_and_trace(pkg, oracle):
    h = pkg.models.random_tb(6, rmax=1, seed=7)
    R, T = h.pack()
    ks = pkg.regulargrid(nk=25)
    rho, eb = oracle.densitymatrix(R, T, ks, R, mu=0.1, Temp=0.05)
    keys = [tuple(R[:, r]) for r in range(R.shape[1])]
    for r, k in enumerate(keys):
        mr = keys.index(tuple(-x for x in k))
        assert np.max(np.abs(rho[:, :, mr] - rho[:, :, r].conj().T)) < 1e-14
    b = oracle.bands(R, T, ks)
    f = oracle.fermidirac(b, 0.1, 0.05)
    i0 = keys.index((0, 0))
    assert abs(np.real(np.trace(rho[:, :, i0])) - f.sum() / ks.shape[1]) < 1e-13
    assert abs(eb - np.sum(b * f) / ks.shape[1]) < 1e-13


def test_threaded_equals_serial(pkg, oracle):
    # test_parallelism.jl:49-106: threaded reduce equals serial to 1e-10
    h = pkg.models.honeycomb_nn(-1.0)
    R, T = h.pack()
    ks = pkg.regulargrid(nk=144)
    r1, e1 = oracle.densitymatrix(R, T, ks, R, 0.0, 0.05, nthreads=1)
    r4, e4 = oracle.densitymatrix(R, T, ks, R, 0.0, 0.05, nthreads=4)
    assert np.max(np.abs(r1 - r4)) < 1e-10 and abs(e1 - e4) < 1e-10


def test_chemicalpotential_rules(oracle):
    bands = np.array([[0.0, 2.0, 4.0], [1.0, 3.0, 5.0]])
    # T = 0: i = 1 + floor(f (L-1)), mu = (E_i + E_{i+1})/2 on the sorted list
    assert oracle.chemicalpotential(bands, 0.5, 0.0) == 2.5
    assert oracle.chemicalpotential(bands, 1.0, 0.0) == 5.0
    assert oracle.chemicalpotential(bands, 0.0, 0.0) == -1.0
    # T > 0: defining equation
    mu = oracle.chemicalpotential(bands, 0.3, 0.2)
    assert abs(oracle.getfilling(bands, mu, 0.2) - 0.3) < 1e-14


def test_regulargrid_order(pkg):
    ks = pkg.regulargrid(nk=9)
    assert ks.shape == (2, 9)
    assert np.allclose(ks[:, 1], [1 / 3, 0]) and np.allclose(ks[:, 3], [0, 1 / 3])
    assert pkg.regulargrid(nk=1000, dim=3).shape == (3, 1000)


def test_random_model_hermitian(pkg):
    h = pkg.models.random_tb(8, rmax=2, seed=2)
    assert len(h) == 25 and h.ishermitian(1e-15)
    k = np.array([0.13, 0.71])
    assert np.max(np.abs(h(k) - h(k).conj().T)) < 1e-14


def test_roots_midpoint_and_bisection(oracle):
    # Roots.jl's Float64 midpoint: bit-pattern midpoint, 0 across a sign change
    assert oracle._middle(1.0, 2.0) == 1.5
    assert oracle._middle(1.0, 4.0) == 2.0
    assert oracle._middle(-4.0, -1.0) == -2.0
    assert oracle._middle(-1.0, 2.0) == 0.0
    # reaches adjacent doubles around an irrational root
    r = oracle._bisect_doubles(lambda x: x * x - 2.0, 0.0, 2.0)
    assert abs(r - np.sqrt(2.0)) <= np.spacing(np.sqrt(2.0))
    # on a plateau of exact zeros the first midpoint inside it is returned
    f = lambda x: 0.0 if 0.3 <= x <= 0.9 else (x - 0.6)
    r = oracle._bisect_doubles(f, -1.0, 2.0)
    assert 0.3 <= r <= 0.9
    # bit midpoints from 0 walk through the exponents: the first point tried above 0.3
    a, b = 0.0, 2.0
    while True:
        m = oracle._middle(a, b)
        if 0.3 <= m <= 0.9:
            break
        a, b = (m, b) if m < 0.3 else (a, m)
    assert r == m
