"""Bogoliubov-de Gennes (Nambu) operators on the GPU.

The reference's fixtures (test/test_superconductivity.jl) run through the GPU: Nambu doubling
and the +-E spectrum, the 2 Delta gap at the Dirac point (1e-10), the construction guards, bond
pairing, the kinetic energy <H_phys> = -3 of graphene at Gamma (:123-143, atol 1e-3), the
filling at mu = 0 with weak pairing (:150-166), the filling on a grid vs nk (:181-198, 1e-12)
and the complex hole block of electron-sector updates (:200-213).  Beyond them, the Nambu
density matrix, the filling and the chemical potential against the FP64 oracle run on the same
Nambu Hops shifted by -mu tau_z: rho to 1e-9, energies to 1e-10, mu through the defining
equation (filling(mu) = target to 1e-12 at T > 0) and to 1e-9 against the oracle's own
bisection.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bdg(pkg):
    import importlib

    return importlib.import_module(pkg.__name__ + ".bdg")


def _graphene(pkg):
    return pkg.models.honeycomb_nn(-1.0)


def _eig(pkg, H, ks):
    b, _ = pkg.api.bandmatrix(H.h, np.asarray(ks, dtype=float).T)
    return b


def _shifted(H, mu):
    """Host copy of the Nambu Hops with -mu tau_z on the zero key (the oracle's input)."""
    h = H.h.copy()
    zk = h.zerokey()
    if zk not in h:
        h[zk] = np.zeros((h.N, h.N), complex)
    d = h.N // 2
    h.data[zk] = h.data[zk] + np.diag(np.concatenate([np.full(d, -mu), np.full(d, mu)]))
    return h


def _oracle_filling(oracle, H, ks, mu, T):
    h = _shifted(H, mu)
    R, Tt = h.pack()
    Rr = np.zeros((h.d, 1), np.int32)
    rho, _ = oracle.densitymatrix(R, Tt, ks, Rr, mu=0.0, Temp=T)
    d = h.N // 2
    return float(np.real(np.trace(rho[:d, :d, 0]))) / d


def test_nambu_doubling_and_pm_spectrum(pkg, bdg, gpu):
    h = _graphene(pkg)
    H = bdg.BdGOperator(h)
    assert H.hopdim == h.N and H.h.N == 2 * h.N
    ev = _eig(pkg, H, [[0.1, 0.2], [0.3, -0.1], [0.0, 0.0]])
    assert np.max(np.abs(ev + ev[::-1])) < 1e-10


def test_swave_gap_at_dirac_point(pkg, bdg, gpu):
    H = bdg.BdGOperator(_graphene(pkg))
    bdg.addpairing(H, pkg.Hops({(0, 0): 0.25 * np.eye(2)}))
    ev = _eig(pkg, H, [[1 / 3, -1 / 3]])[:, 0]
    n = len(ev)
    assert np.max(np.abs(ev + ev[::-1])) < 1e-10
    assert abs(ev[n // 2] - 0.25) < 1e-10 and abs(ev[n // 2 - 1] + 0.25) < 1e-10


def test_construction_guards(pkg, bdg):
    bad = pkg.Hops({(1, 0): [[0, 1.0], [0, 0]], (0, 0): np.zeros((2, 2))})
    with pytest.raises(AssertionError):
        bdg.BdGOperator(bad)
    H = bdg.BdGOperator(_graphene(pkg))
    with pytest.raises(RuntimeError):
        bdg.addpairing(H, pkg.Hops({(1, 0): 0.1 * np.eye(2)}))
    H = bdg.BdGOperator(_graphene(pkg))
    bdg.addpairing(H, pkg.Hops({(1, 0): [[0.1, 0.2], [0, 0.1]], (-1, 0): [[0.1, 0], [0, 0.1]]}))
    assert H.ishermitian()


def test_bond_pairing_hermitian_pm_spectrum(pkg, bdg, gpu):
    bond = np.array([[0, 0.05], [0, 0]], complex)
    H = bdg.BdGOperator(_graphene(pkg), pkg.Hops({(1, 0): bond, (-1, 0): bond.conj().T}))
    assert H.ishermitian()
    ks = [[0.1, 0.2], [0.3, -0.1], [1 / 3, -1 / 3]]
    for k in ks:
        Hk = H(k)
        assert np.max(np.abs(Hk - Hk.conj().T)) < 1e-12
    ev = _eig(pkg, H, ks)
    assert np.max(np.abs(ev + ev[::-1])) < 1e-10


def test_kinetic_energy_graphene_gamma(pkg, bdg, gpu):
    # test_superconductivity.jl:123-143: graphene at Gamma, mu = 0, T -> 0: <H> = -3
    H = bdg.BdGOperator(_graphene(pkg))
    rho = bdg.BdGOperator(pkg.Hops({R: np.zeros((2, 2), complex) for R in _graphene(pkg).keys()}))
    ek = bdg.getdensitymatrix(rho, H, np.zeros((2, 1)), 0.0, T=1e-6)
    assert abs(ek - (-3.0)) < 1e-3
    assert abs(ek - (-3.0)) < 1e-12  # the occupations are exactly 0/1 at T = 1e-6 here


def test_filling_weak_pairing(pkg, bdg, oracle, gpu):
    H = bdg.BdGOperator(_graphene(pkg))
    bdg.addpairing(H, pkg.Hops({(0, 0): 0.05 * np.eye(2)}))
    ks = pkg.regulargrid(nk=12 ** 2)
    f0 = bdg.filling(H, ks, 0.0, T=0.05)
    assert abs(f0 - 0.5) < 5e-3
    assert abs(f0 - _oracle_filling(oracle, H, ks, 0.0, 0.05)) < 1e-12
    # filling on an explicit grid equals the nk form
    Hp = bdg.BdGOperator(_graphene(pkg), pkg.Hops({(0, 0): 0.2 * np.eye(2)}))
    fm = bdg.filling(Hp, pkg.regulargrid(nk=8 ** 2), 0.1, T=0.05)
    fn = bdg.filling(Hp, None, 0.1, T=0.05, nk=8)
    assert abs(fm - fn) < 1e-12
    assert abs(fm - _oracle_filling(oracle, Hp, pkg.regulargrid(nk=64), 0.1, 0.05)) < 1e-12


def test_electron_projector_and_chemical_potential_blocks(pkg, bdg):
    H = bdg.BdGOperator(pkg.Hops({(0, 0): np.diag([1.0, -1.0])}))
    P = bdg.electron(H)((0.0, 0.0))
    assert np.allclose(P, np.diag([1, 1, 0, 0])) and np.allclose(P @ P, P)
    h = _graphene(pkg)
    Hb = bdg.addchemicalpotential(bdg.BdGOperator(h), 0.2)
    Hk = Hb((0.0, 0.0))
    assert np.allclose(Hk[:2, :2], h((0, 0)) + 0.2 * np.eye(2))
    assert np.allclose(Hk[2:, 2:], -np.conj(h((0, 0))) - 0.2 * np.eye(2))
    assert Hb.ishermitian()


def test_electron_sector_update_keeps_complex_hole_block(pkg, bdg, gpu):
    h0 = pkg.Hops({(0,): np.zeros((1, 1))})
    dh = pkg.Hops({(1,): [[1.0 + 0.3j]], (-1,): [[1.0 - 0.3j]]})
    H = bdg.addelectronsector(bdg.BdGOperator(h0), dh)
    full = h0.copy()
    for R, t in dh.items():
        full[R] = t
    E = bdg.BdGOperator(full)
    for k in ([0.0], [0.17], [0.41]):
        assert np.allclose(H(k), E(k))
    b1, _ = pkg.api.bandmatrix(H.h, np.array([[0.0, 0.17, 0.41]]))
    b2, _ = pkg.api.bandmatrix(E.h, np.array([[0.0, 0.17, 0.41]]))
    assert np.max(np.abs(b1 - b2)) < 1e-14


def _random_nambu(pkg, bdg, n, seed, delta):
    h = pkg.models.addspin(pkg.models.random_tb(n, rmax=1, seed=seed), "spinhalf")
    # singlet s-wave pairing Delta (x) i sigma_y on every site
    isy = np.array([[0, 1], [-1, 0]], complex)
    return bdg.BdGOperator(h, pkg.Hops({(0, 0): np.kron(delta * np.eye(n), isy)}))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n,nk", [(4, 36), (32, 16)])
def test_nambu_densitymatrix_vs_oracle(pkg, bdg, oracle, n, nk, precision):
    # 8 and 64 spinful orbitals -> Nambu 16 (per-k kernels) and 128 (the large-N layout)
    H = _random_nambu(pkg, bdg, n, 70 + n, 0.3)
    ks = pkg.regulargrid(nk=nk)
    mu, T = 0.15, 0.03
    keys = [(0, 0), (1, 0), (-1, 1)]
    rho = bdg.BdGOperator.from_nambu(pkg.Hops({R: np.zeros((H.h.N, H.h.N), complex) for R in keys}))
    ex = pkg.api.to_executor("gpu", precision=precision)
    ek = bdg.getdensitymatrix(rho, H, ks, mu, T=T, executor=ex)
    hs = _shifted(H, mu)
    R, Tt = hs.pack()
    Rr = np.array(keys, np.int32).T
    ref, eb = oracle.densitymatrix(R, Tt, ks, Rr, mu=0.0, Temp=T)
    for r, key in enumerate(keys):
        assert np.max(np.abs(rho[key] - ref[:, :, r])) < 1e-9, key
    d = H.hopdim
    ek_ref = eb / 2 + (np.real(np.trace(H.h[(0, 0)][:d, :d])) - d * mu) / 2
    assert abs(ek - ek_ref) < 1e-10 * max(1.0, abs(ek_ref))
    # the model on the device is restored: a second call gives the same result bit for bit
    rho2 = rho.zero()
    assert bdg.getdensitymatrix(rho2, H, ks, mu, T=T, executor=ex) == ek


def test_nambu_chemical_potential_vs_oracle(pkg, bdg, oracle, gpu):
    H = _random_nambu(pkg, bdg, 6, 91, 0.4)
    ks = pkg.regulargrid(nk=25)
    T, target = 0.05, 0.37
    mu, passes = bdg.chemicalpotential(H, ks, target, T=T, return_passes=True)
    assert 3 <= passes <= 70
    assert abs(_oracle_filling(oracle, H, ks, mu, T) - target) < 1e-12
    from oracle.oracle import _bisect_doubles

    bands, _ = pkg.api.bandmatrix(bdg.getelectronsector(H), ks)
    lo, hi = float(bands.min()) - 1.0, float(bands.max()) + 1.0
    mu_ref = _bisect_doubles(lambda x: _oracle_filling(oracle, H, ks, x, T) - target, lo, hi)
    assert abs(mu - mu_ref) < 1e-9
    with pytest.raises(ValueError):
        bdg.chemicalpotential(H, ks, 1.5, T=T)  # not bracketed
