"""The band-topology oracle (oracle/berry.py) pinned on the CPU: the reference's Haldane
fixtures (test/test_haldane.jl: gap 3 sqrt(3) t2 at K, band Chern numbers +-1 flipping with
phi) and closed forms (the Su-Schrieffer-Heeger Zak phase, 1/2 in the topological phase and 0
in the trivial one; gauge invariance of the plaquette phases)."""
import importlib

import numpy as np
import pytest


@pytest.fixture(scope="module")
def lat(pkg):
    return importlib.import_module(pkg.__name__ + ".lattice")


def _haldane(lat, phi):
    L = lat.honeycomb()
    return lat.addhops(lat.nearestneighbor(L, -1.0), lat.haldane(L, 0.1, phi))


def test_haldane_fixtures(lat):
    from oracle import berry as ob

    for phi, s in ((np.pi / 2, 1.0), (-np.pi / 2, -1.0)):
        H = _haldane(lat, phi)
        ev = np.linalg.eigvalsh(H([1 / 3, -1 / 3]))
        assert abs(ev[0] + 0.5196152422706632) < 1e-12 and abs(ev[1] - 0.5196152422706632) < 1e-12
        R, T = H.pack()
        c = ob.getcherns(R, T, 30, 30)
        assert abs(c[0] - s) < 0.05 and abs(c[1] + s) < 0.05


@pytest.mark.parametrize("v,w,zak", [(0.5, 1.0, 0.5), (1.0, 0.5, 0.0)])
def test_ssh_zak_phase(pkg, v, w, zak):
    from oracle import berry as ob
    from oracle import oracle

    # intra-cell hopping v (A-B), inter-cell w (B of cell R to A of cell R + 1)
    h = pkg.Hops({(0,): [[0, v], [v, 0]], (1,): [[0, 0], [w, 0]], (-1,): [[0, w], [0, 0]]})
    R, T = h.pack()
    ks = np.linspace(0.0, 1.0, 101)[:-1].reshape(1, -1)
    _, V = oracle.bands(R, T, np.asfortranarray(ks), vectors=True)
    ph = ob.wilson1d(np.transpose(V[:, [0], :], (2, 0, 1)))[0]
    assert abs(np.exp(2j * np.pi * ph) - np.exp(2j * np.pi * zak)) < 1e-9


def test_plaquettes_gauge_invariant(pkg):
    from oracle import berry as ob

    h = pkg.models.random_tb(4, rmax=1, seed=17)
    R, T = h.pack()
    st = ob.statesgrid(R, T, 6, 5, [0, 1])
    rng = np.random.default_rng(1)
    # an arbitrary unitary rotation of the two-band frame at every grid point
    g = st.copy()
    for i in range(6):
        for j in range(5):
            q, _ = np.linalg.qr(rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2)))
            g[i, j] = st[i, j] @ q
    d = np.angle(np.exp(2j * np.pi * (ob.berry(g) - ob.berry(st))))
    assert np.max(np.abs(d)) < 1e-12
