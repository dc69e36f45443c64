"""The Hartree-Fock oracle and the lattice generators, pinned on the CPU.

* generators: the generic gethops restatement reproduces the hand-written honeycomb
  model and the Hubbard interaction of the SCF configuration;
* HF functional: the variational identity Tr[H_MF rho] - Tr[h rho] = 2 (eps_H + eps_F)
  for a complex rho (test_meanfield.jl:228-258), including an interaction that reaches
  beyond the hopping (test_meanfield.jl:260-278); a Fock term with conj(rho) fails it;
* SCF: Neel anti-alignment at U = 4 (test_meanfield.jl:205-219), the U -> 0 limit
  (test_meanfield.jl:33-54) and the energy identities at the fixed point
  (selfconsistent.jl:36-50).
"""
import numpy as np
import pytest


def _trprod(A, B, ks):
    return sum(np.real(np.trace(A(k) @ B(k))) for k in ks.T) / ks.shape[1]


@pytest.fixture(scope="module")
def hf():
    from oracle import hartreefock

    return hartreefock


@pytest.fixture(scope="module")
def lat(pkg):
    import importlib

    return importlib.import_module(pkg.__name__ + ".lattice")


@pytest.fixture(scope="module")
def mfm(pkg):
    import importlib

    return importlib.import_module(pkg.__name__ + ".meanfield")


def test_generic_hops_match_honeycomb_model(pkg, lat):
    L = lat.honeycomb()
    h = lat.nearestneighbor(L, -1.0)
    ref = pkg.models.honeycomb_nn(-1.0)
    assert set(h.keys()) == set(ref.keys())
    for R in ref.keys():
        assert np.array_equal(h[R], ref[R])
    g = lat.graphene(L)
    # t_graphene: NN -1, next-nearest -exp(-(sqrt3 - 1)/0.08), third shell below tmin
    assert np.isclose(g[(0, 0)][0, 1], -1.0)
    nnn = -np.exp(-(np.sqrt(3) - 1) / 0.08)
    assert np.isclose(g[(1, 0)][0, 0], nnn) and np.isclose(g[(1, 0)][1, 1], nnn)
    assert g.ishermitian()


def test_hubbard_generator(pkg, lat):
    L = lat.honeycomb()
    v = lat.gethubbard(L, mode="σx", a=0.5, U=3.0)
    _, vref, _ = pkg.models.honeycomb_hubbard(3.0)
    assert list(v.keys()) == [(0, 0)]
    assert np.array_equal(v[(0, 0)], vref[(0, 0)])


def _complex_rho(N):
    # test_meanfield.jl:237-241, 1-based i, j
    M = np.array([[np.exp(1j * (3 * i + j)) / (1 + abs(i - j)) for j in range(1, N + 1)]
                  for i in range(1, N + 1)])
    return (M + M.conj().T) / 2


def test_variational_identity_complex_rho(pkg, lat, hf):
    L = lat.honeycomb()
    h = lat.graphene(L, mode="spinhalf")
    v = lat.gethubbard(L, mode="σx", a=0.5, U=4.0)
    rho = {(0, 0): _complex_rho(4)}
    assert np.max(np.abs(rho[(0, 0)].imag)) > 0.1
    hmf, eH, eF = hf.meanfield(h.data, v.data, rho)
    H = pkg.Hops(hmf)
    R = pkg.Hops(rho)
    ks = np.array([[i / 3, j / 3] for i in range(3) for j in range(3)]).T
    lhs = _trprod(H, R, ks) - _trprod(h, R, ks)
    assert abs(lhs - 2 * (eH + eF)) < 1e-10
    assert H.ishermitian()
    # the conjugated Fock field of the old reference code fails the identity
    Hc = pkg.Hops({K: b.copy() for K, b in hmf.items()})
    Hc[(0, 0)] = Hc[(0, 0)] + v[(0, 0)] * rho[(0, 0)] - v[(0, 0)] * rho[(0, 0)].conj()
    bad = _trprod(Hc, R, ks) - _trprod(h, R, ks)
    assert abs(bad - 2 * (eH + eF)) > 1e-3


def test_variational_identity_extended_interaction(pkg, lat, hf, mfm):
    L = lat.honeycomb()
    h = pkg.Hops({(0, 0): np.diag([0.3, 0.3, -0.3, -0.3]).astype(complex)})
    nbrs = [(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, -1), (1, -1), (-1, 1)]
    v = lat.gethubbard(L, nbrs, mode="σx", a=1.8, U=2.0)
    assert not set(v.keys()) <= set(h.keys())
    rho = mfm.initialguess(v, "antiferro", sublattice=L.label("sublattice"))
    hmf, eH, eF = hf.meanfield(h.data, v.data, rho.data)
    assert np.isfinite(eH) and np.isfinite(eF)
    H = pkg.Hops(hmf)
    assert H.ishermitian()
    ks = np.array([[i / 3, j / 3] for i in range(3) for j in range(3)]).T
    lhs = _trprod(H, rho, ks) - _trprod(h, rho, ks)
    assert abs(lhs - 2 * (eH + eF)) < 1e-10


def _sublattice_mags(mfm, rho, sub):
    PA = mfm.sublatticeprojection(sub, 0)
    PB = mfm.sublatticeprojection(sub, 1)
    return mfm.magnetization(rho, PA), mfm.magnetization(rho, PB)


def test_oracle_scf_neel_order(pkg, lat, hf, mfm, oracle):
    L = lat.honeycomb()
    h = lat.graphene(L, mode="spinhalf")
    v = lat.gethubbard(L, mode="σx", a=0.5, U=4.0)
    rho0 = mfm.initialguess(v, "antiferro", sublattice=L.label("sublattice"))
    out = hf.scf(h.data, v.data, rho0.data, 0.5, pkg.regulargrid(nk=30 ** 2), T=0.01, beta=0.2,
                 tol=1e-5, iterations=800)
    assert out["converged"]
    a, b = _sublattice_mags(mfm, pkg.Hops(out["rho"]), L.label("sublattice"))
    assert abs(np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b) + 1.0) < 1e-8
    assert np.linalg.norm(a) > 0.1  # ordered at U = 4


def test_oracle_scf_small_U_reproduces_bands(pkg, lat, hf, mfm, oracle):
    L = lat.honeycomb()
    h = pkg.models.addspin(lat.nearestneighbor(L, -1.0), "spinhalf")
    v = lat.gethubbard(L, mode="σx", a=0.5, U=1e-6)
    rho0 = mfm.initialguess(v, "antiferro", sublattice=L.label("sublattice"))
    out = hf.scf(h.data, v.data, rho0.data, 0.5, pkg.regulargrid(nk=30 ** 2), T=0.0, beta=0.7,
                 tol=1e-5, iterations=800)
    assert out["converged"]
    R, T = hf._pack(out["hmf"])
    ks = np.random.default_rng(3).random((2, 100))
    b = oracle.bands(R, T, ks) - out["mu"]
    rs = [(1, 0), (-1, 0), (0, 1), (0, -1), (1, -1), (-1, 1)]
    up = np.sqrt(3 + sum(np.cos(2 * np.pi * (ks[0] * r[0] + ks[1] * r[1])) for r in rs))
    for n, sgn in enumerate([-1, -1, 1, 1]):
        assert np.mean(np.abs(b[n] - sgn * up)) < 1e-5


def test_oracle_scf_energy_identities(pkg, lat, hf, mfm, oracle):
    h, v, sub = pkg.models.honeycomb_hubbard(3.0)
    rho0 = mfm.staggered_guess(v, sub, m=0.1, noise=1e-3, seed=1)
    for R in h.keys():  # rho on keys(h): Tr[h rho] is then exact
        if R not in rho0:
            rho0[R] = np.zeros((4, 4), complex)
    out = hf.scf(h.data, v.data, rho0.data, 0.5, pkg.regulargrid(nk=24 ** 2), T=0.01, beta=1.0,
                 tol=1e-10, iterations=200)
    assert out["converged"]
    rho = pkg.Hops(out["rho"])
    kin = np.real(mfm.trace(rho, h))
    assert abs(out["ekin"] - kin) < 1e-8
    assert abs(out["E"] - (out["ekin"] + out["eH"] + out["eF"])) < 1e-12
    assert abs(out["E"] - 0.5 * (out["ekin"] + out["eband"])) < 1e-8
    # filling: tr rho(0) = filling * N
    assert abs(np.real(np.trace(rho[(0, 0)])) - 2.0) < 1e-8
