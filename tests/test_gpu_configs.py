"""The benchmark configurations at their own sizes, GPU vs the FP64 CPU oracle.

Bars (the stated parity targets): eigenvalues |e - e_ref| <= 1e-10 max(|e_ref|, ||H(k)||_2)
(the relative bar, made well-posed at e = 0), rho(R) to 1e-9 absolute, E_band to 1e-10
relative, SCF order parameters (rho of the converged mean field) and energies to 1e-8,
dOmega/dt_R against the oracle's rho(-R)^T to 1e-9 and against central differences
(h = 1e-6, FP64) to 1e-6 relative.

* random_tb N=64, 25 keys, 128 x 128 grid, mu = 0, kT = 0.02: the headline workload, FP64 and
  mixed (the grid is one chunk, so the compact-WY T factors run on the side stream);
* the same model on 48 x 48 (2304 k-points: still the side-stream fork) against the oracle;
* random_tb N=128, 9 keys, seed 5, 64 x 64 grid, mu = 0, kT = 0.05: Omega and its gradient;
* models drawn like the 256-model batch (N = even round(exp U(ln 16, ln 512)), 25 keys, Hubbard
  U ~ U(1, 4) in the sigma_x form, seed 4): the first 16 draws with N <= 200 -- nine models,
  three of them above 112 orbitals (large-N path in both precisions) -- on 8 x 8 grids, batch SCF with the
  FP64 and the mixed executor.
"""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mfm(pkg):
    return importlib.import_module(pkg.__name__ + ".meanfield")


@pytest.fixture(scope="module")
def hf():
    from oracle import hartreefock

    return hartreefock


def _hnorm(h, ks):
    return np.array([np.linalg.norm(h(k), 2) for k in ks.T])


def _threads():
    import os

    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return 1


def _rho_check(pkg, oracle, h, ks, mu, T, ex, tol=1e-9):
    R, Tt = h.pack()
    ref, eb_ref = oracle.densitymatrix(R, Tt, ks, R, mu=mu, Temp=T, nthreads=_threads())
    rho = pkg.Hops({key: np.zeros((h.N, h.N), complex) for key in h.keys()})
    eb = pkg.api.getdensitymatrix(rho, h, ks, mu, T=T, executor=ex)
    err = max(np.max(np.abs(rho[tuple(R[:, r])] - ref[:, :, r])) for r in range(R.shape[1]))
    assert err < tol, err
    assert abs(eb - eb_ref) <= 1e-10 * max(1.0, abs(eb_ref)), (eb, eb_ref)
    return err


@pytest.fixture(scope="module")
def headline(pkg):
    h = pkg.models.random_tb(64, rmax=2, seed=2)
    assert len(h.keys()) == 25
    return h


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_headline_full_grid_vs_oracle(pkg, oracle, gpu, headline, precision):
    h = headline
    ks = pkg.regulargrid(nk=128 * 128)
    ex = pkg.api.to_executor("gpu", precision=precision)
    _rho_check(pkg, oracle, h, ks, 0.0, 0.02, ex)
    assert ex.last_precision() == precision
    if precision == "mixed":
        rep = ex.refinement_report()
        assert rep["sweeps"] >= 1 and rep["kpoints_rho_fused"] > 0
    R, Tt = h.pack()
    ref = oracle.bands(R, Tt, ks, nthreads=_threads())
    b, _ = pkg.api.bandmatrix(h, ks, executor=ex)
    err = np.abs(b - ref) / np.maximum(np.abs(ref), _hnorm(h, ks)[None, :])
    assert np.max(err) < 1e-10, np.max(err)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_side_stream_tfactors_vs_oracle(pkg, oracle, gpu, headline, precision):
    # 2304 k-points: the T factors of the compact WY run on the context's side stream, forked
    # after the tridiagonalization and joined before the back-transformation
    ex = pkg.api.to_executor("gpu", precision=precision)
    err = _rho_check(pkg, oracle, headline, pkg.regulargrid(nk=48 * 48), 0.0, 0.02, ex,
                     tol=1e-10 if precision == "fp64" else 1e-9)
    assert err >= 0.0


def test_grandpotential_adjoint_full_size(pkg, oracle, gpu):
    N, mu, T = 128, 0.0, 0.05
    h = pkg.models.random_tb(N, rmax=1, seed=5)
    assert len(h.keys()) == 9
    ks = pkg.regulargrid(nk=64 * 64)
    om, g = pkg.api.grandpotential(h, ks, mu, T, gradient=True)
    R, Tt = h.pack()
    b = oracle.bands(R, Tt, ks, nthreads=_threads())
    x = (b - mu) / T
    om_ref = float(np.mean(np.sum(-T * (np.maximum(-x, 0) + np.log1p(np.exp(-np.abs(x)))), axis=0)))
    assert abs(om - om_ref) <= 1e-12 * max(1.0, abs(om_ref))
    rho, _ = oracle.densitymatrix(R, Tt, ks, -R, mu=mu, Temp=T, nthreads=_threads())
    for r, key in enumerate(h.keys()):
        assert np.max(np.abs(g[key] - rho[:, :, r].T)) < 1e-9, key
    # central differences of the GPU Omega on sampled entries, under the Hermitian tie
    # t_{-R} = t_R^H: d/ds Omega(t_R + s E_ij, t_-R + s E_ji) = 2 Re g_R[i, j] for R != 0 and,
    # for a real symmetric onsite perturbation E_ij + E_ji, g_0[i, j] + g_0[j, i]
    _, pull = pkg.api.grandpotential_rrule(h, ks, mu, T)
    cot = pull(1.0)
    rng = np.random.default_rng(55)
    eps = 1e-6
    for key in [(1, 0), (1, -1), (0, 1), (0, 0)]:
        i, j = (int(v) for v in rng.integers(0, N, 2))
        mkey = tuple(-v for v in key)
        D = np.zeros((N, N), complex)
        D[i, j] = 1.0 + 0.5j if key != mkey else 1.0
        if key == mkey:
            D = D + D.conj().T

        def omega(s):
            hp = h.copy()
            hp[key] = h[key] + s * D
            if key != mkey:
                hp[mkey] = h[mkey] + s * D.conj().T
            return pkg.api.grandpotential(hp, ks, mu, T)

        fd = (omega(eps) - omega(-eps)) / (2 * eps)
        pred = np.real(np.sum(np.conj(cot[key]) * D))
        if key != mkey:
            pred += np.real(np.sum(np.conj(cot[mkey]) * D.conj().T))
        assert abs(pred - fd) <= 1e-6 * max(1.0, abs(fd)), (key, i, j, pred, fd)


def _batch_models(pkg, mfm, nmax=200, draws=16):
    """Models drawn like the 256-model batch (seed 4), keeping the first draws with N <= nmax."""
    rng = np.random.default_rng(4)
    sx = np.array([[0, 1], [1, 0]], dtype=complex)
    ks = pkg.regulargrid(nk=8 * 8)
    probs = []
    for m in range(draws):
        N = int(2 * round(np.exp(rng.uniform(np.log(16), np.log(512))) / 2))
        U = float(rng.uniform(1, 4))
        if N > nmax:
            continue
        h = pkg.models.random_tb(N, rmax=2, seed=4000 + m)
        v = pkg.Hops({(0, 0): U * np.kron(np.eye(N // 2), sx)})
        rho0 = mfm.staggered_guess(v, np.arange(N // 2) % 2, m=0.2, noise=1e-3, seed=m)
        probs.append((h, v, rho0, 0.5, ks))
    return probs


@pytest.fixture(scope="module")
def batch_ref(pkg, mfm, hf):
    probs = _batch_models(pkg, mfm)
    kw = dict(T=0.02, beta=0.5, tol=1e-9, iterations=400)
    refs = [hf.scf(h.data, v.data, rho0.data, fill, ks, nthreads=_threads(), **kw)
            for (h, v, rho0, fill, ks) in probs]
    return probs, kw, refs


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_uneven_batch_scf_vs_oracle(pkg, mfm, batch_ref, precision):
    probs, kw, refs = batch_ref
    Ns = [p[0].N for p in probs]
    assert len(probs) == 9 and sum(N > 112 for N in Ns) >= 3 and all(len(p[0].keys()) == 25 for p in probs)
    ex = pkg.api.to_executor("gpu", precision=precision)
    out = mfm.solvehartreefock_batch(probs, nstreams=4, executor=ex, **kw)
    for N, ref, (rho, E, info, conv, res) in zip(Ns, refs, out):
        assert ref["converged"] and conv, (N, res)
        assert abs(E - ref["E"]) < 1e-8, (N, E, ref["E"])
        for R in ref["rho"]:
            assert np.max(np.abs(rho[R] - ref["rho"][R])) < 1e-8, (N, R)


def test_honeycomb_hubbard_outputs_vs_oracle(pkg, oracle, gpu, mfm, hf):
    # the converged mean field of the spinful honeycomb Hubbard SCF (U = 3, half filling,
    # kT = 0.01, 60 x 60 grid, beta = 0.5, tol 1e-8): its DOS on 801 frequencies over [-4, 4]
    # measured from mu (Lorentzian of width 0.02 without 1/pi, the reference's getdos kernel)
    # and its bands on the path Gamma-K-M-Gamma, both on the GPU, against the oracle on the
    # same H_MF; the mean field itself against the oracle SCF's
    h, v, sub = pkg.models.honeycomb_hubbard(3.0)
    rho0 = mfm.staggered_guess(v, sub, m=0.1, noise=1e-3, seed=1)
    ks = pkg.regulargrid(nk=60 ** 2)
    kw = dict(T=0.01, beta=0.5, tol=1e-8, iterations=500)
    rho, E, mf, conv, res = mfm.solvehartreefock(h, v, rho0, 0.5, ks, executor=gpu, **kw)
    ref = hf.scf(h.data, v.data, rho0.data, 0.5, ks, **kw)
    assert conv and ref["converged"]
    for R, blk in ref["hmf"].items():
        assert np.max(np.abs(mf.hMF[R] - blk)) < 1e-8, R
    R, Tt = mf.hMF.pack()
    om = np.linspace(-4.0, 4.0, 801) + mf.mu
    dos = pkg.api.getdos(mf.hMF, om, ks, Gamma=0.02, executor=gpu)
    dref = oracle.dos(R, Tt, ks, om, 0.02, nthreads=_threads())
    assert np.max(np.abs(dos - dref)) <= 1e-10 * np.max(dref)
    path = pkg.kpath(np.array([[0, 0], [1 / 3, -1 / 3], [0.5, 0], [0, 0]]).T, num_points=301)
    b, _ = pkg.api.bandmatrix(mf.hMF, path, executor=gpu)
    bref = oracle.bands(R, Tt, path)
    err = np.abs(b - bref) / np.maximum(np.abs(bref), _hnorm(mf.hMF, path)[None, :])
    assert np.max(err) < 1e-10
    # the antiferromagnet is gapped at half filling: mu lies between bands 2 and 3
    assert np.max(b[1]) < mf.mu < np.min(b[2])
