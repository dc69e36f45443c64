"""The eigenvector-consumer oracle (oracle/response.py) and the current operators pinned on the
CPU, by routes that do not share its algebra:

* Kubo: the closed form of a two-level model with non-Hermitian currents (which orders and
  conjugates the matrix elements: M_12 = a conj(c), M_21 = b conj(d)); the resolvent form
  sum_m f_m u_m^H J1 (e_m - z - H)^-1 J2^H u_m + sum_n f_n u_n^H J2^H (e_n + z - H)^-1 J1 u_n
  (linear solves, no pair list) on a random model; the reference's graphene invariants
  (test/test_linearresponse.jl: sigma_xx = sigma_yy, Re sigma_xy = 0, Re sigma_xx >= 0) and
  the universal interband conductivity of the Dirac cone, sigma_0 = A_c / 8 per unit cell
  (e = hbar = 1, no spin), within 2 % at w = 0.4 t on a fine grid;
* ldos: the resolvent -Im[(w + iG - H(k))^-1]_ii / pi, and sum_i ldos_i = getdos / pi at one
  frequency (the reference's normalisation note, dos.jl:96-101);
* density: the diagonal of the T = 0 density matrix rho(R = 0) of the C oracle;
* getcurrentoperators: Hermitian J(k) (test/test_operators.jl:135-152) and J_a(k) =
  d/ds sum_L t_L e^{2 pi i k.L} e^{-i s (r_i + A L - r_j)_a} at s = 0 by central differences.
"""
import importlib

import numpy as np
import pytest


@pytest.fixture(scope="module")
def lat(pkg):
    return importlib.import_module(pkg.__name__ + ".lattice")


@pytest.fixture(scope="module")
def orr(oracle):
    from oracle import response

    return response


def test_kubo_two_level_closed_form(pkg, orr):
    e1, e2 = -0.7, 0.4
    a, b, c, d = 0.3 + 0.2j, -0.5 + 0.1j, 0.8 - 0.4j, 0.25 + 0.6j
    H = pkg.Hops({(0,): np.diag([e1, e2])})
    J1 = pkg.Hops({(0,): [[0, a], [b, 0]]})
    J2 = pkg.Hops({(0,): [[0, c], [d, 0]]})
    ks = np.zeros((1, 1))
    ws = np.array([0.0, 0.3, 1.1, 2.5])
    G, T, mu = 0.07, 0.2, 0.1
    R, Tt = H.pack()
    R1, T1 = J1.pack()
    R2, T2 = J2.pack()
    chi = orr.kubo(R, Tt, R1, T1, R2, T2, ks, ws, mu, G, T)
    f1, f2 = 1 / (np.exp((e1 - mu) / T) + 1), 1 / (np.exp((e2 - mu) / T) + 1)
    z = ws + 1j * G
    ref = -1j * ((f1 - f2) * a * np.conj(c) / (e1 - e2 - z) + (f2 - f1) * b * np.conj(d) / (e2 - e1 - z))
    assert np.max(np.abs(chi - ref)) < 1e-14


def test_kubo_resolvent_form(pkg, orr, oracle):
    h = pkg.models.random_tb(6, rmax=1, seed=41)
    rng = np.random.default_rng(43)
    keys = list(h.keys())
    J1 = pkg.Hops({R: rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6)) for R in keys})
    J2 = pkg.Hops({R: rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6)) for R in keys[:3]})
    ks = pkg.regulargrid(nk=9)
    ws = np.array([0.0, 0.4, 1.7])
    G, T, mu = 0.1, 0.15, 0.05
    R, Tt = h.pack()
    chi = orr.kubo(R, Tt, *J1.pack(), *J2.pack(), ks, ws, mu, G, T)
    ref = np.zeros(len(ws), complex)
    for k in ks.T:
        Hk = h(k)
        e, U = np.linalg.eigh(Hk)
        A, B = J1(k), J2(k).conj().T
        f = oracle.fermidirac(e, mu, T)
        I = np.eye(6)
        for o, w in enumerate(ws):
            z = w + 1j * G
            for m in range(6):
                u = U[:, m]
                ref[o] += f[m] * (u.conj() @ A @ np.linalg.solve((e[m] - z) * I - Hk, B @ u))
                ref[o] += f[m] * (u.conj() @ B @ np.linalg.solve((e[m] + z) * I - Hk, A @ u))
    ref *= -1j / ks.shape[1]
    assert np.max(np.abs(chi - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


def _sigma(chi, ws, G):
    return -(chi - chi[0]) / (ws + 1j * G)


def test_graphene_reference_invariants(pkg, lat, orr):
    # test/test_linearresponse.jl: graphene (mode nospin), klin = 30, 30 frequencies in
    # [0.05, 3], G = 0.05, T = 0.01, mu = 0
    L = lat.honeycomb()
    H = lat.graphene(L)
    J = lat.getcurrentoperators(L, H)
    assert len(J) == 2
    ks = pkg.regulargrid(nk=30 ** 2)
    ws = np.linspace(0.05, 3.0, 30)
    R, Tt = H.pack()
    chi = {ab: orr.kubo(R, Tt, *J[ab[0]].pack(), *J[ab[1]].pack(), ks, ws, 0.0, 0.05, 0.01)
           for ab in ((0, 0), (1, 1), (0, 1))}
    sxx, syy, sxy = (_sigma(chi[ab], ws, 0.05) for ab in ((0, 0), (1, 1), (0, 1)))
    assert np.max(np.abs(sxx.real - syy.real)) < 1e-10
    assert np.max(np.abs(sxy.real)) < 1e-10
    assert np.all(sxx.real > -1e-6)


def test_graphene_universal_conductivity(pkg, lat, orr):
    L = lat.honeycomb()
    H = lat.nearestneighbor(L, -1.0)
    J = lat.getcurrentoperators(L, H)
    ks = pkg.regulargrid(nk=180 ** 2)
    ws = np.array([0.0, 0.4])
    R, Tt = H.pack()
    s = _sigma(orr.kubo(R, Tt, *J[0].pack(), *J[0].pack(), ks, ws, 0.0, 0.05, 0.01), ws, 0.05)
    Ac = abs(np.linalg.det(L.A[:2, :2]))
    assert abs(s[1].real / (Ac / 8) - 1.0) < 0.02


def test_ldos_resolvent_and_dos_sum(pkg, orr, oracle):
    h = pkg.models.random_tb(5, rmax=1, seed=7)
    ks = pkg.regulargrid(nk=16)
    R, Tt = h.pack()
    G = 0.2
    for w in (-0.3, 0.9):
        n = orr.ldos(R, Tt, ks, [w], G)
        ref = np.zeros(5)
        for k in ks.T:
            ref += -np.imag(np.diag(np.linalg.inv((w + 1j * G) * np.eye(5) - h(k)))) / np.pi
        ref /= ks.shape[1]
        assert np.max(np.abs(n - ref)) < 1e-13
        assert abs(np.sum(n) - oracle.dos(R, Tt, ks, np.array([w]), G)[0] / np.pi) < 1e-13
    # a window is the mean of its single-frequency results
    ws = [-0.3, 0.2, 0.9]
    mean = sum(orr.ldos(R, Tt, ks, [w], G) for w in ws) / 3
    assert np.max(np.abs(orr.ldos(R, Tt, ks, ws, G) - mean)) < 1e-14


def test_density_is_rho_diagonal(pkg, orr, oracle):
    h = pkg.models.random_tb(6, rmax=1, seed=8)
    ks = pkg.regulargrid(nk=25)
    R, Tt = h.pack()
    mu = 0.1
    b = oracle.bands(R, Tt, ks)
    assert np.min(np.abs(b - mu)) > 1e-6  # no state at mu: the step and heaviside agree
    rho, _ = oracle.densitymatrix(R, Tt, ks, np.zeros((2, 1), np.int32), mu=mu, Temp=0.0)
    n = orr.density(R, Tt, ks, mu)
    assert np.max(np.abs(n - np.real(np.diag(rho[:, :, 0])))) < 1e-13
    assert abs(np.sum(n) - np.sum(b <= mu) / ks.shape[1]) < 1e-12


@pytest.mark.parametrize("spin", [False, True])
def test_current_operators_hermitian_and_derivative(pkg, lat, spin):
    L = lat.honeycomb()
    H = lat.nearestneighbor(L, -1.0)
    H = lat.addhops(H, lat.haldane(L, 0.1, 0.3))
    if spin:
        H = pkg.models.addspin(H, "spinhalf")
    Js = lat.getcurrentoperators(L, H)
    assert len(Js) == 2
    n = H.N // L.n
    pos = np.repeat(L.positions, n, axis=1)
    for k in ([0.0, 0.0], [1 / 3, -1 / 3], [0.5, 0.5], [-0.2, 0.7]):
        for a, J in enumerate(Js):
            Jk = J(k)
            assert np.max(np.abs(Jk - Jk.conj().T)) < 1e-10

            def F(s):
                out = np.zeros((H.N, H.N), complex)
                for Lk, t in H.items():
                    d = (pos[a][:, None] + (L.A @ np.asarray(Lk, float))[a] - pos[a][None, :])
                    out += t * np.exp(2j * np.pi * np.dot(k, Lk)) * np.exp(-1j * s * d)
                return out

            eps = 1e-5
            fd = (F(eps) - F(-eps)) / (2 * eps)
            assert np.max(np.abs(fd - Jk)) < 1e-8


def test_current_operators_from_pair_function(pkg, lat):
    L = lat.honeycomb()

    def t(r1, r2):
        r = np.linalg.norm(r1[:3] - r2[:3])
        return -1.0 if 0.99 < r < 1.01 else 0.0

    Jf = lat.getcurrentoperators(L, t, cellrange=2)
    Jh = lat.getcurrentoperators(L, lat.nearestneighbor(L, -1.0))
    for k in ([0.1, 0.3], [1 / 3, -1 / 3]):
        for a in range(2):
            assert np.max(np.abs(Jf[a](k) - Jh[a](k))) < 1e-12
