"""Grand potential and spectral adjoints vs the FP64 oracle and finite differences.

Omega to 1e-12 relative (oracle eigenvalues, closed-form sum); dOmega/dt_R = rho(-R)^T
to 1e-9 against the oracle density matrix; directional derivatives against central
differences of the oracle's Omega (h = 1e-6, FP64) to 1e-6 relative, the gradient bar.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _omega_ref(oracle, h, ks, mu, T):
    R, Tt = h.pack()
    b = oracle.bands(R, Tt, ks)
    x = (b - mu) / T
    return float(np.mean(np.sum(-T * (np.maximum(-x, 0) + np.log1p(np.exp(-np.abs(x)))), axis=0)))


@pytest.mark.parametrize("N,rmax,nk,mu,T", [(8, 1, 64, 0.0, 0.05), (24, 1, 36, 0.3, 0.1), (64, 1, 16, 0.0, 0.05)])
def test_grandpotential_and_gradient(pkg, oracle, gpu, N, rmax, nk, mu, T):
    h = pkg.models.random_tb(N, rmax=rmax, seed=500 + N)
    ks = pkg.regulargrid(nk=nk)
    om, g = pkg.api.grandpotential(h, ks, mu, T, gradient=True)
    ref = _omega_ref(oracle, h, ks, mu, T)
    assert abs(om - ref) <= 1e-12 * max(1.0, abs(ref))
    R, Tt = h.pack()
    Rneg = -R
    rho, _ = oracle.densitymatrix(R, Tt, ks, Rneg, mu=mu, Temp=T)
    for r, key in enumerate(h.keys()):
        assert np.max(np.abs(g[key] - rho[:, :, r].T)) < 1e-9


def test_gradient_vs_finite_differences(pkg, oracle, gpu):
    h = pkg.models.random_tb(16, rmax=1, seed=77)
    ks = pkg.regulargrid(nk=64)
    mu, T, eps = 0.1, 0.05, 1e-6
    om, pull = pkg.api.grandpotential_rrule(h, ks, mu, T)
    cot = pull(1.0)
    rng = np.random.default_rng(5)
    for key in [(0, 0), (1, 0), (1, -1)]:
        D = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
        mkey = tuple(-x for x in key)
        if key == mkey:
            D = (D + D.conj().T) / 2

        def shifted(s):
            hp = h.copy()
            hp[key] = h[key] + s * D
            if key != mkey:
                hp[mkey] = h[mkey] + s * D.conj().T
            return hp

        fd = (_omega_ref(oracle, shifted(eps), ks, mu, T) - _omega_ref(oracle, shifted(-eps), ks, mu, T)) / (2 * eps)
        # cotangent convention: dOmega = Re sum conj(cot_R) D over every perturbed block
        pred = np.real(np.sum(np.conj(cot[key]) * D))
        if key != mkey:
            pred += np.real(np.sum(np.conj(cot[mkey]) * D.conj().T))
        assert abs(pred - fd) <= 1e-6 * max(1.0, abs(fd)), (key, pred, fd)


def test_grandpotential_T0_and_weights(pkg, oracle, gpu):
    h = pkg.models.random_tb(6, rmax=1, seed=9)
    ks = pkg.regulargrid(nk=49)
    R, Tt = h.pack()
    b = oracle.bands(R, Tt, ks)
    om = pkg.api.grandpotential(h, ks, 0.2, 0.0)
    assert abs(om - float(np.mean(np.sum(np.minimum(b - 0.2, 0.0), axis=0)))) < 1e-12
    w = np.random.default_rng(1).random(49)
    omw = pkg.api.grandpotential(h, ks, 0.2, 0.05, kweights=w)
    x = (b - 0.2) / 0.05
    ref = float(np.sum(w * np.sum(-0.05 * (np.maximum(-x, 0) + np.log1p(np.exp(-np.abs(x)))), axis=0)))
    assert abs(omw - ref) < 1e-12 * max(1, abs(ref))


def test_weighted_densitymatrix_and_bandloss(pkg, oracle, gpu):
    h = pkg.models.random_tb(10, rmax=1, seed=21)
    ks = pkg.regulargrid(nk=25)
    R, Tt = h.pack()
    b, V = oracle.bands(R, Tt, ks, vectors=True)
    occ = np.random.default_rng(2).standard_normal(b.shape)
    rhos = pkg.Hops({key: np.zeros((10, 10), complex) for key in [(0, 0), (1, 0), (2, -1)]})
    pkg.api.densitymatrix_occ(rhos, h, ks, occ)
    for key in rhos.keys():
        ref = sum(np.exp(-2j * np.pi * np.dot(ks[:, j], key)) * (V[:, :, j] * occ[:, j]) @ V[:, :, j].conj().T
                  for j in range(ks.shape[1])) / ks.shape[1]
        assert np.max(np.abs(rhos[key] - ref)) < 1e-10
    # band-loss cotangent vs central differences along a Hermitian-tied direction
    target = b + 0.1 * np.random.default_rng(3).standard_normal(b.shape)
    L, cot = pkg.api.bandloss(h, ks, target)
    assert abs(L - np.mean(np.sum((b - target) ** 2, axis=0))) < 1e-12
    D = np.random.default_rng(4).standard_normal((10, 10)) + 0j

    def loss(s):
        hp = h.copy()
        hp[(1, 0)] = h[(1, 0)] + s * D
        hp[(-1, 0)] = h[(-1, 0)] + s * D.T
        Rp, Tp = hp.pack()
        return np.mean(np.sum((oracle.bands(Rp, Tp, ks) - target) ** 2, axis=0))

    fd = (loss(1e-6) - loss(-1e-6)) / 2e-6
    pred = np.real(np.sum(np.conj(cot[(1, 0)]) * D) + np.sum(np.conj(cot[(-1, 0)]) * D.T))
    assert abs(pred - fd) <= 1e-6 * max(1, abs(fd))


def test_bandfit_descends(pkg, gpu):
    # fit a model to the bands of a second seeded model with plain gradient steps
    ks = pkg.regulargrid(nk=64)
    target, _ = pkg.api.bandmatrix(pkg.models.random_tb(8, rmax=1, seed=2), ks)
    h = pkg.models.random_tb(8, rmax=1, seed=1)
    L0, _ = pkg.api.bandloss(h, ks, target)
    lr = 0.1
    losses = []
    for _ in range(40):
        L, cot = pkg.api.bandloss(h, ks, target)
        losses.append(L)
        for key in h.keys():
            h[key] = h[key] - lr * cot[key]
    assert losses[-1] < 0.1 * L0
    assert all(b < a for a, b in zip(losses, losses[1:]))
    assert h.ishermitian(1e-12)


def _mf_problem(pkg, n=6, seed=31):
    """Spinful random model with onsite Hubbard (sigma_x form) plus a nearest-cell interaction
    (v_{-L} = v_L^T), and a Hermitian density matrix on the keys of v."""
    rng = np.random.default_rng(seed)
    h = pkg.models.addspin(pkg.models.random_tb(n, rmax=1, seed=seed), "spinhalf")
    N = h.N
    sx = np.array([[0, 1], [1, 0]], dtype=complex)
    vl = 0.3 * rng.standard_normal((N, N))
    v = pkg.Hops({(0, 0): 2.5 * np.kron(np.eye(n), sx), (1, 0): vl, (-1, 0): vl.T})
    X = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    r1 = 0.1 * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    rho = pkg.Hops({(0, 0): (X + X.conj().T) / 4 + 0.5 * np.eye(N), (1, 0): r1, (-1, 0): r1.conj().T})
    return h, v, rho


def test_meanfield_grandpotential_and_parameter_gradients(pkg, oracle, gpu):
    import importlib

    mfm = importlib.import_module(pkg.__name__ + ".meanfield")
    from oracle import hartreefock as hfo

    h, v, rho = _mf_problem(pkg)
    ks = pkg.regulargrid(nk=36)
    mu, T = 0.1, 0.05
    mf = mfm.HartreeFock(h, v, executor=gpu)
    om, dh, dv = mf.grandpotential(rho, ks, mu, T, gradient=True)
    # Omega of the oracle's mean field, and dOmega/dh_R = rho_MF(-R)^T
    hmf, _, _ = hfo.meanfield(h.data, v.data, rho.data)
    H = pkg.Hops(hmf)
    assert abs(om - _omega_ref(oracle, H, ks, mu, T)) <= 1e-12 * max(1.0, abs(om))
    R, Tt = H.pack()
    Rh = np.array(list(h.keys()), np.int32).T
    rneg, _ = oracle.densitymatrix(R, Tt, ks, -Rh, mu=mu, Temp=T)
    for r, key in enumerate(h.keys()):
        assert np.max(np.abs(dh[key] - rneg[:, :, r].T)) < 1e-9, key

    # dOmega/dv against central differences of the oracle's Omega(H_MF(rho; v + s D))
    def omega_v(vp):
        hm, _, _ = hfo.meanfield(h.data, vp.data, rho.data)
        return _omega_ref(oracle, pkg.Hops(hm), ks, mu, T)

    rng = np.random.default_rng(8)
    eps = 1e-6
    N = h.N
    for key in [(0, 0), (1, 0)]:
        D = rng.standard_normal((N, N))
        if key == (0, 0):
            D = D + D.T
        mkey = (-key[0], -key[1])

        def shifted(s):
            vp = v.copy()
            vp[key] = v[key] + s * D
            if key != mkey:
                vp[mkey] = v[mkey] + s * D.T
            return vp

        fd = (omega_v(shifted(eps)) - omega_v(shifted(-eps))) / (2 * eps)
        pred = np.sum(dv[key] * D) + (np.sum(dv[mkey] * D.T) if key != mkey else 0.0)
        assert abs(pred.imag) < 1e-9 * max(1.0, abs(pred))
        assert abs(pred.real - fd) <= 1e-6 * max(1.0, abs(fd)), (key, pred, fd)
    # the Hubbard strength: v(0) = U (I (x) sigma_x)
    P = np.kron(np.eye(N // 2), np.array([[0, 1], [1, 0]]))
    dU = np.real(np.sum(dv[(0, 0)] * P))
    vU = lambda s: pkg.Hops({**{k: b for k, b in v.items()}, (0, 0): v[(0, 0)] + s * P})
    fdU = (omega_v(vU(eps)) - omega_v(vU(-eps))) / (2 * eps)
    assert abs(dU - fdU) <= 1e-6 * max(1.0, abs(fdU))
    # the pullback's cotangents are the conjugated derivatives
    om2, pull = mf.grandpotential_rrule(rho, ks, mu, T)
    ch, cv = pull(2.0)
    assert om2 == om and np.allclose(cv[(1, 0)], 2.0 * np.conj(dv[(1, 0)]), rtol=0, atol=0)
    # the GPU Omega is the function the gradient describes: its own central difference
    fdg = (mfm.HartreeFock(h, vU(eps), executor=gpu).grandpotential(rho, ks, mu, T)
           - mfm.HartreeFock(h, vU(-eps), executor=gpu).grandpotential(rho, ks, mu, T)) / (2 * eps)
    assert abs(dU - fdg) <= 1e-6 * max(1.0, abs(fdg))


def test_bandfit_128_orbitals_200_steps(pkg, gpu):
    # fit the 9 hopping blocks of a 128-orbital model (seed 5) to the bands of a second seeded
    # model (seed 6) on the 64 x 64 grid: 200 plain gradient steps through the band-loss pullback
    ks = pkg.regulargrid(nk=64 * 64)
    target, _ = pkg.api.bandmatrix(pkg.models.random_tb(128, rmax=1, seed=6), ks)
    h = pkg.models.random_tb(128, rmax=1, seed=5)
    lr = 0.3
    losses = []
    for _ in range(200):
        L, cot = pkg.api.bandloss(h, ks, target)
        losses.append(L)
        for key in h.keys():
            h[key] = h[key] - lr * cot[key]
    assert all(b <= a * (1 + 1e-12) for a, b in zip(losses, losses[1:]))
    assert losses[-1] < 0.1 * losses[0], (losses[0], losses[-1])
    assert h.ishermitian(1e-10)
