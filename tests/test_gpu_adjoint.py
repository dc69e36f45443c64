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
