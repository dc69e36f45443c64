"""Band topology on the GPU: Berry plaquette phases, Chern numbers, gap Chern numbers, Wilson
loops and the path curvature, against the reference's Haldane fixtures (test/test_haldane.jl)
and the FP64 oracle (oracle/berry.py, LAPACK eigenvectors).

Bars: the Haldane gap at K to 1e-12 (the reference's pinned value); band Chern numbers to 0.05
and gap Chern numbers exactly (the reference's bars); plaquette phases, Chern sums, Wilson
eigenphases (as e^{2 pi i phase}) and path curvatures to 1e-9 against the oracle -- they are
gauge invariant, so the GPU and LAPACK eigenvector phases need not agree; determinant phases of
the link kernel to 1e-12 against numpy.
"""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def berry(pkg):
    return importlib.import_module(pkg.__name__ + ".berry")


@pytest.fixture(scope="module")
def lat(pkg):
    return importlib.import_module(pkg.__name__ + ".lattice")


@pytest.fixture(scope="module")
def ob():
    from oracle import berry as b

    return b


def _haldane(lat, phi):
    L = lat.honeycomb()
    return lat.addhops(lat.nearestneighbor(L, -1.0), lat.haldane(L, 0.1, phi))


def test_haldane_gap_at_K(pkg, lat, gpu):
    b, _ = pkg.api.bandmatrix(_haldane(lat, np.pi / 2), np.array([[1 / 3], [-1 / 3]]))
    assert abs(b[0, 0] + 0.5196152422706632) < 1e-12 and abs(b[1, 0] - 0.5196152422706632) < 1e-12


@pytest.mark.parametrize("phi,sign", [(np.pi / 2, 1), (-np.pi / 2, -1)])
def test_haldane_band_and_gap_cherns(lat, berry, gpu, phi, sign):
    H = _haldane(lat, phi)
    c = berry.getcherns(H, 30, 30)
    assert abs(c[0] - sign) < 0.05 and abs(c[1] + sign) < 0.05
    gaps = berry.gapcherns(H, 30, gaptol=1e-3)
    assert len(gaps) == 1
    n, lo, hi, ch = gaps[0]
    assert n == 1 and ch == sign and lo < 0 < hi


def _random4(pkg):
    return pkg.models.random_tb(4, rmax=1, seed=17)


def test_berry_plaquettes_vs_oracle(pkg, berry, ob, gpu):
    # non-abelian flux of the two lowest bands of a random 4-orbital model on a 9 x 7 grid
    h = _random4(pkg)
    R, T = h.pack()
    _, _, st = berry.statesgrid(h, 9, 7, [0, 1])
    F = berry.berry(st)
    Fr = ob.berry(ob.statesgrid(R, T, 9, 7, [0, 1]))
    assert F.shape == (8, 6)
    d = np.angle(np.exp(2j * np.pi * (F - Fr))) / (2 * np.pi)
    assert np.max(np.abs(d)) < 1e-9
    mid, F2 = berry.berry(h, 9, 7, [0, 1])
    assert np.array_equal(F2, F) and mid.shape == (8, 6, 2)


def test_getcherns_vs_oracle(pkg, berry, ob, gpu):
    h = pkg.models.random_tb(6, rmax=1, seed=23)
    R, T = h.pack()
    c = berry.getcherns(h, 12, 10)
    cr = ob.getcherns(R, T, 12, 10)
    assert np.max(np.abs(np.array(c) - np.array(cr))) < 1e-9
    assert abs(berry.getwindnum(h, 12, 10, 2) - sum(cr[:3])) < 1e-9


def test_wilson_loop_vs_oracle(pkg, berry, ob, gpu):
    # a one-dimensional 4-orbital model: Wilson loop of its two lowest bands
    rng = np.random.default_rng(3)
    t0 = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    t1 = 0.7 * (rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))
    h = pkg.Hops({(0,): (t0 + t0.conj().T) / 2, (1,): t1, (-1,): t1.conj().T})
    en, U, st = berry.wilson1d(h, 25, [0, 1])
    ks = np.linspace(0.0, 1.0, 25).reshape(1, -1)
    R, T = h.pack()
    from oracle import oracle

    _, V = oracle.bands(R, T, np.asfortranarray(ks), vectors=True)
    ref = ob.wilson1d(np.transpose(V[:, [0, 1], :], (2, 0, 1)))
    z, zr = np.sort_complex(np.exp(2j * np.pi * en)), np.sort_complex(np.exp(2j * np.pi * ref))
    assert np.max(np.abs(z - zr)) < 1e-9
    assert U.shape == (2, 2) and st.shape == (25, 4, 2)


def test_berryalongpath_vs_numpy(pkg, berry, lat, gpu):
    H = _haldane(lat, np.pi / 2)
    path = pkg.kpath(np.array([[0, 0], [1 / 3, -1 / 3], [0.5, 0]]).T, num_points=21)
    bc = berry.berryalongpath(H, path)
    assert bc.shape == (2, path.shape[1])
    # the same plaquettes from numpy eigenvectors
    kp = np.column_stack([path[:, 0] - (path[:, 1] - path[:, 0]), path, path[:, -1] + (path[:, -1] - path[:, -2])])
    kp = kp + 1.34e-6 * np.linalg.norm(kp[:, 1] - kp[:, 0]) * np.array([1.0, 0.7071067811865476])[:, None]
    rot = lambda t: np.array([[np.cos(t), -np.sin(t)], [np.sin(t), np.cos(t)]])
    ref = np.zeros_like(bc)
    for j in range(path.shape[1]):
        kc = kp[:, j + 1]
        dR, dL = (kp[:, j + 2] - kc) / 2, (kp[:, j] - kc) / 2
        dU = (rot(np.pi / 2) @ dR + rot(-np.pi / 2) @ dL) / 2
        dD = (rot(-np.pi / 2) @ dR + rot(np.pi / 2) @ dL) / 2
        Us = [np.linalg.eigh(H(kc + dk))[1] for dk in (dR, dU, dD, dL)]
        for n in range(2):
            Lk = lambda a, b: (lambda z: z / abs(z))(np.vdot(Us[a][:, n], Us[b][:, n]))
            ref[n, j] = np.angle(Lk(0, 1) * Lk(1, 3) / Lk(2, 3) / Lk(0, 2)) / (2 * np.pi)
    assert np.max(np.abs(bc - ref)) < 1e-9


@pytest.mark.parametrize("M", [1, 3, 17, 40])
def test_link_phase_kernel(berry, gpu, M):
    rng = np.random.default_rng(M)
    N, npts = 40, 6
    blocks = rng.standard_normal((npts, N, M)) + 1j * rng.standard_normal((npts, N, M))
    pairs = [(0, 1), (2, 3), (5, 4), (1, 1), (3, 0)]
    ph = berry.link_phases(blocks, pairs)
    O = berry.link_overlaps(blocks, pairs)
    for l, (a, b) in enumerate(pairs):
        ref = blocks[a].conj().T @ blocks[b]
        assert np.max(np.abs(O[l] - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
        d = np.linalg.det(ref)
        assert abs(ph[l] - d / abs(d)) < 1e-12
    # an exactly singular overlap (a zero state set) gives the (0, 0) phase
    sing = blocks.copy()
    sing[0] = 0.0
    assert berry.link_phases(sing, [(0, 1)])[0] == 0
