"""Eigenvector consumers on the GPU: ldos, density and the Kubo current-current correlator,
against the FP64 oracle (oracle/response.py) on the same inputs, and the reference's own
optical-conductivity invariants (test/test_linearresponse.jl) on the GPU result.

Bars: ldos and density to 1e-12 relative to the largest entry; chi(w) to 1e-10 relative to
max |chi| (normwise over the frequencies: the pair sums cancel between occupied and empty
states, so single entries carry the rounding of the largest terms); the reference's
invariants with its own tolerances.  Paths covered: per-k eigensolver (N <= 64, FP64 and the
mixed executor, whose eigenvectors come from the FP64 path), the large-N blocked eigensolver
(N = 160), several frequency tiles (Nw > 64), k-weights, the 128 x 128 headline grid in two
chunks of k-points (Kubo at N = 64), and J1 == J2 (one U^H J U product).
"""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def resp(pkg):
    return importlib.import_module(pkg.__name__ + ".response")


@pytest.fixture(scope="module")
def lat(pkg):
    return importlib.import_module(pkg.__name__ + ".lattice")


@pytest.fixture(scope="module")
def orr(oracle):
    from oracle import response

    return response


def _rand_hops(pkg, keys, N, seed):
    rng = np.random.default_rng(seed)
    return pkg.Hops({R: rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)) for R in keys})


@pytest.mark.parametrize("N,nk,precision", [(6, 25, "fp64"), (40, 64, "fp64"), (40, 64, "mixed"), (160, 9, "fp64")])
def test_ldos_density_vs_oracle(pkg, orr, N, nk, precision):
    h = pkg.models.random_tb(N, rmax=1, seed=100 + N)
    ks = pkg.regulargrid(nk=nk)
    R, T = h.pack()
    ex = pkg.api.to_executor("gpu", precision=precision)
    for ws in ([0.3], [-1.0, 0.0, 0.7, 2.0]):
        n = pkg.api.ldos(h, ks, ws, Gamma=0.15, executor=ex)
        ref = orr.ldos(R, T, ks, ws, 0.15)
        assert np.max(np.abs(n - ref)) < 1e-12 * np.max(np.abs(ref)), ws
    d = pkg.api.density(h, ks, 0.2, executor=ex)
    dref = orr.density(R, T, ks, 0.2)
    assert np.max(np.abs(d - dref)) < 1e-12 * max(1.0, np.max(dref))
    # k-weights replace the 1/Nk
    w = np.linspace(0.5, 1.5, ks.shape[1])
    n = pkg.api.ldos(h, ks, [0.3], Gamma=0.15, kweights=w, executor=ex)
    ref = orr.ldos(R, T, ks, [0.3], 0.15, kweights=w)
    assert np.max(np.abs(n - ref)) < 1e-12 * np.max(np.abs(ref))


def test_ldos_sums_to_dos(pkg, gpu):
    h = pkg.models.random_tb(16, rmax=1, seed=3)
    ks = pkg.regulargrid(nk=100)
    n = pkg.api.ldos(h, ks, 0.4, Γ=0.1, executor=gpu)
    dos = pkg.api.getdos(h, np.array([0.4]), ks, Gamma=0.1, executor=gpu)
    assert abs(np.sum(n) - dos[0] / np.pi) < 1e-12 * dos[0]


def _sigma(chi, ws, G):
    return -(chi - chi[0]) / (ws + 1j * G)


def test_graphene_optical_conductivity_reference_invariants(pkg, lat, resp, orr, gpu):
    # test/test_linearresponse.jl on the GPU: klin = 30, 30 frequencies, G = 0.05, T = 0.01
    L = lat.honeycomb()
    H = lat.graphene(L)
    ws = np.linspace(0.05, 3.0, 30)
    kw = dict(klin=30, mu=0.0, Gamma=0.05, T=0.01, executor=gpu)
    cxx = resp.opticalconductivity(ws, 0, 0, H, L, **kw)
    cyy = resp.opticalconductivity(ws, 1, 1, H, L, **kw)
    cxy = resp.opticalconductivityXY(ws, H, L, **kw)
    sxx, syy, sxy = (_sigma(c, ws, 0.05) for c in (cxx, cyy, cxy))
    assert np.max(np.abs(sxx.real - syy.real)) < 1e-10
    assert np.max(np.abs(sxy.real)) < 1e-10
    assert np.all(sxx.real > -1e-6)
    J = lat.getcurrentoperators(L, H)
    R, T = H.pack()
    ks = pkg.regulargrid(nk=900)
    # chi_xy vanishes by symmetry (its entries are rounding noise): both on the scale of chi_xx
    scale = None
    for c, (a, b) in ((cxx, (0, 0)), (cxy, (0, 1))):
        ref = orr.kubo(R, T, *J[a].pack(), *J[b].pack(), ks, ws, 0.0, 0.05, 0.01)
        scale = scale or np.max(np.abs(ref))
        assert np.max(np.abs(c - ref)) < 1e-10 * scale
    assert np.max(np.abs(resp.opticalconductivityXX(ws, H, L, **kw) - cxx)) == 0.0


@pytest.mark.parametrize("N,nk,Nw,precision,same", [
    (8, 36, 20, "fp64", False),
    (24, 49, 150, "fp64", False),   # three frequency tiles
    (24, 49, 150, "mixed", True),
    (160, 4, 70, "fp64", False),    # the large-N eigensolver
])
def test_kubo_vs_oracle(pkg, resp, orr, N, nk, Nw, precision, same):
    h = pkg.models.random_tb(N, rmax=1, seed=200 + N)
    keys = list(h.keys())
    J1 = _rand_hops(pkg, keys, N, 1)
    J2 = J1 if same else _rand_hops(pkg, keys[:4], N, 2)
    ks = pkg.regulargrid(nk=nk)
    ws = np.linspace(0.0, 4.0, Nw)
    ex = pkg.api.to_executor("gpu", precision=precision)
    for T in (0.0, 0.05):
        chi = resp.opticalconductivity(ws, h, J1, J2, ks, mu=0.1, Gamma=0.05, T=T, executor=ex)
        R, Tt = h.pack()
        ref = orr.kubo(R, Tt, *J1.pack(), *J2.pack(), ks, ws, 0.1, 0.05, T)
        assert np.max(np.abs(chi - ref)) < 1e-10 * np.max(np.abs(ref)), T


def test_kubo_kweights(pkg, resp, orr, gpu):
    h = pkg.models.random_tb(12, rmax=1, seed=9)
    J = _rand_hops(pkg, list(h.keys()), 12, 5)
    ks = pkg.regulargrid(nk=36)
    w = np.linspace(0.2, 2.0, ks.shape[1])
    ws = np.linspace(0.0, 2.0, 11)
    chi = resp.opticalconductivity(ws, h, J, J, ks, kweights=w, Gamma=0.1, T=0.02, executor=gpu)
    R, T = h.pack()
    ref = orr.kubo(R, T, *J.pack(), *J.pack(), ks, ws, 0.0, 0.1, 0.02, kweights=w)
    assert np.max(np.abs(chi - ref)) < 1e-10 * np.max(np.abs(ref))


def test_kubo_headline_grid_two_chunks(pkg, resp, orr, gpu):
    # N = 64 with 25 keys on the 128 x 128 grid: five N x N buffers per k-point put the grid in
    # two chunks of k-points, the partial sums carried across them
    h = pkg.models.random_tb(64, rmax=2, seed=2)
    J = _rand_hops(pkg, list(h.keys()), 64, 11)
    J2 = _rand_hops(pkg, list(h.keys())[:5], 64, 12)
    ks = pkg.regulargrid(nk=128 * 128)
    ws = np.linspace(0.0, 3.0, 16)
    chi = resp.opticalconductivity(ws, h, J, J2, ks, mu=0.0, Gamma=0.05, T=0.02, executor=gpu)
    R, T = h.pack()
    ref = orr.kubo(R, T, *J.pack(), *J2.pack(), ks, ws, 0.0, 0.05, 0.02)
    assert np.max(np.abs(chi - ref)) < 1e-10 * np.max(np.abs(ref))


def test_argument_errors(pkg, resp, gpu):
    h = pkg.models.random_tb(4, rmax=1, seed=1)
    J = pkg.models.random_tb(6, rmax=1, seed=2)
    with pytest.raises(Exception):
        resp.opticalconductivity([0.1], h, J, J, pkg.regulargrid(nk=4), executor=gpu)
    with pytest.raises(Exception):
        pkg.api.ldos(h, pkg.regulargrid(nk=4), [], executor=gpu)


def test_edge_cases(pkg, resp, orr, gpu):
    # one orbital: no pairs (chi = 0), ldos a single Lorentzian, density 0 / 1 around the band
    h = pkg.Hops({(0,): [[0.2]], (1,): [[0.5]], (-1,): [[0.5]]})
    ks = np.linspace(0.0, 1.0, 8, endpoint=False).reshape(1, -1)
    chi = resp.opticalconductivity([0.0, 0.5], h, h, h, ks, Gamma=0.1, T=0.05, executor=gpu)
    assert np.max(np.abs(chi)) == 0.0
    e = 0.2 + 2 * 0.5 * np.cos(2 * np.pi * ks[0])
    n = pkg.api.ldos(h, ks, 0.3, Gamma=0.1, executor=gpu)
    assert abs(n[0] - np.mean(0.1 / ((0.3 - e) ** 2 + 0.01)) / np.pi) < 1e-13
    assert pkg.api.density(h, ks, -5.0, executor=gpu)[0] == 0.0
    assert abs(pkg.api.density(h, ks, 5.0, executor=gpu)[0] - 1.0) < 1e-15
    # T = 0 with mu exactly on an eigenvalue: a diagonal model (eigenvalues -1, 0, 1 exact in
    # every implementation) gives the step with 1/2 at e = mu (fermidirac) in the Kubo
    # occupations and e <= mu in the density, as the oracle
    h2 = pkg.Hops({(0,): np.diag([-1.0, 0.0, 1.0])})
    J = pkg.Hops({(0,): [[0, 0.3, 0.1j], [0.3, 0, 0.2], [-0.1j, 0.2, 0]]})
    R, T = h2.pack()
    ks2 = np.zeros((1, 1))
    ws = np.array([0.0, 0.4, 1.3])
    chi = resp.opticalconductivity(ws, h2, J, J, ks2, mu=0.0, Gamma=0.1, T=0.0, executor=gpu)
    ref = orr.kubo(R, T, *J.pack(), *J.pack(), ks2, ws, 0.0, 0.1, 0.0)
    assert np.max(np.abs(chi - ref)) < 1e-13 * np.max(np.abs(ref))
    # closed form with f = (1, 1/2, 0): pairs (0, 1), (1, 2), (0, 2)
    z = ws + 1j * 0.1
    Jm = np.array([[0, 0.3, 0.1j], [0.3, 0, 0.2], [-0.1j, 0.2, 0]])
    e, f = np.array([-1.0, 0.0, 1.0]), np.array([1.0, 0.5, 0.0])
    cf = sum((f[i] - f[j]) * abs(Jm[i, j]) ** 2 / (e[i] - e[j] - z) for i in range(3) for j in range(3) if i != j)
    assert np.max(np.abs(chi - (-1j) * cf)) < 1e-13
    d = pkg.api.density(h2, ks2, 0.0, executor=gpu)
    assert np.array_equal(d, [1.0, 1.0, 0.0])
