"""GPU Hartree-Fock functional and self-consistent loop vs the FP64 oracle.

Tolerances: H_MF blocks 1e-13 absolute and eps_H / eps_F 1e-12 (elementwise FP64, same
operation order); SCF order parameters, energies, mu and rho 1e-8 (the order-parameter
bar); iteration counts equal (same residual sequence to ~1e-13 of tol); the variational
identity 1e-10 (test_meanfield.jl:228-258).
"""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mfm(pkg):
    return importlib.import_module(pkg.__name__ + ".meanfield")


@pytest.fixture(scope="module")
def lat(pkg):
    return importlib.import_module(pkg.__name__ + ".lattice")


@pytest.fixture(scope="module")
def hf():
    from oracle import hartreefock

    return hartreefock


def _random_herm_hops(pkg, keys, N, rng, scale=1.0):
    out = pkg.Hops()
    for R in keys:
        mR = tuple(-x for x in R)
        if R in out:
            continue
        X = scale * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
        if R == mR:
            out[R] = (X + X.conj().T) / 2
        else:
            out[R] = X
            out[mR] = X.conj().T
    return out


@pytest.mark.parametrize("flags", [(True, True), (True, False), (False, True)])
def test_meanfield_vs_oracle(pkg, gpu, mfm, hf, flags):
    rng = np.random.default_rng(11)
    N = 6
    h = pkg.models.random_tb(N, rmax=1, seed=4)
    vkeys = [(0, 0), (1, 0), (-1, 0), (2, 1), (-2, -1)]  # (2, 1) is beyond the hopping range
    v = _random_herm_hops(pkg, vkeys, N, rng, 0.5)
    rho = _random_herm_hops(pkg, [(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1), (2, 1), (-2, -1)], N, rng)
    mf = mfm.HartreeFock(h, v, hartree=flags[0], fock=flags[1], executor=gpu)(rho)
    ref, eH, eF = hf.meanfield(h.data, v.data, rho.data, hartree=flags[0], fock=flags[1])
    assert list(mf.hMF.keys()) == list(ref.keys())
    for R in ref:
        assert np.max(np.abs(mf.hMF[R] - ref[R])) < 1e-13
    assert abs(mf.epsH - eH) <= 1e-12 * max(1, abs(eH))
    assert abs(mf.epsF - eF) <= 1e-12 * max(1, abs(eF))


def test_variational_identity_on_gpu(pkg, gpu, mfm, lat):
    L = lat.honeycomb()
    h = lat.graphene(L, mode="spinhalf")
    v = lat.gethubbard(L, mode="σx", a=0.5, U=4.0)
    M = np.array([[np.exp(1j * (3 * i + j)) / (1 + abs(i - j)) for j in range(1, 5)] for i in range(1, 5)])
    rho = pkg.Hops({(0, 0): (M + M.conj().T) / 2})
    mf = mfm.HartreeFock(h, v, executor=gpu)(rho)
    ks = np.array([[i / 3, j / 3] for i in range(3) for j in range(3)]).T
    tr = lambda A, B: sum(np.real(np.trace(A(k) @ B(k))) for k in ks.T) / ks.shape[1]
    assert abs(tr(mf.hMF, rho) - tr(h, rho) - 2 * mf.doublecounting()) < 1e-10
    assert mf.hMF.ishermitian()


def _sub_mags(mfm, rho, sub):
    return (mfm.magnetization(rho, mfm.sublatticeprojection(sub, 0)),
            mfm.magnetization(rho, mfm.sublatticeprojection(sub, 1)))


def _compare_scf(pkg, mfm, gpu_out, ref, sub, tol=1e-8, T=0.0, ks=None, filling=0.5):
    rho, E, mf, conv, res = gpu_out
    assert conv == ref["converged"]
    assert mf.history.shape[0] == ref["iterations"]
    assert abs(E - ref["E"]) < tol
    if T == 0.0:
        assert abs(mf.mu - ref["mu"]) < tol
    else:
        # In a gap getfilling(mu) = filling holds exactly on a whole interval, so the
        # bisection root depends on the summation order of the filling; check that the
        # GPU mu solves the defining equation on the oracle's bands instead.
        from oracle import hartreefock, oracle

        R, Tm = hartreefock._pack(ref["hmf"])
        b = oracle.bands(R, Tm, ks)
        assert abs(oracle.getfilling(b, mf.mu, T) - filling) < 1e-10
    assert abs(mf.epsband - ref["eband"]) < tol and abs(mf.epskin - ref["ekin"]) < tol
    assert abs(mf.epsH - ref["eH"]) < tol and abs(mf.epsF - ref["eF"]) < tol
    for R in ref["rho"]:
        assert np.max(np.abs(rho[R] - ref["rho"][R])) < tol
    ga, gb = _sub_mags(mfm, rho, sub)
    ra, rb = _sub_mags(mfm, pkg.Hops(ref["rho"]), sub)
    assert np.max(np.abs(ga - ra)) < tol and np.max(np.abs(gb - rb)) < tol
    # the per-iteration energies follow the same trajectory
    assert np.max(np.abs(mf.history[:, 0] - ref["history"][:, 0])) < 1e-7


def test_scf_honeycomb_hubbard_60x60_vs_oracle(pkg, gpu, mfm, hf):
    # spinful honeycomb Hubbard, U = 3, half filling, kT = 0.01, 60 x 60 grid, beta = 0.5,
    # tol 1e-8, staggered m = 0.1 start with 1e-3 Hermitian noise (seed 1)
    h, v, sub = pkg.models.honeycomb_hubbard(3.0)
    rho0 = mfm.staggered_guess(v, sub, m=0.1, noise=1e-3, seed=1)
    ks = pkg.regulargrid(nk=60 ** 2)
    out = mfm.solvehartreefock(h, v, rho0, 0.5, ks, T=0.01, beta=0.5, tol=1e-8, iterations=500,
                               executor=gpu)
    ref = hf.scf(h.data, v.data, rho0.data, 0.5, ks, T=0.01, beta=0.5, tol=1e-8, iterations=500)
    assert ref["converged"]
    _compare_scf(pkg, mfm, out, ref, sub, T=0.01, ks=ks)
    a, b = _sub_mags(mfm, out[0], sub)
    assert abs(np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b) + 1) < 1e-8


def test_scf_T0_and_extra_keys_vs_oracle(pkg, gpu, mfm, hf):
    h, v, sub = pkg.models.honeycomb_hubbard(3.5)
    rho0 = mfm.staggered_guess(v, sub, m=0.2, noise=1e-3, seed=2)
    for R in h.keys():
        if R not in rho0:
            rho0[R] = np.zeros((4, 4), complex)
    ks = pkg.regulargrid(nk=21 ** 2)
    out = mfm.solvehartreefock(h, v, rho0, 0.5, ks, T=0.0, beta=0.6, tol=1e-9, iterations=300,
                               executor=gpu)
    ref = hf.scf(h.data, v.data, rho0.data, 0.5, ks, T=0.0, beta=0.6, tol=1e-9, iterations=300)
    _compare_scf(pkg, mfm, out, ref, sub)
    # energy bookkeeping at the fixed point
    rho, E, mf = out[0], out[1], out[2]
    assert abs(mf.epskin - np.real(mfm.trace(rho, h))) < 1e-7
    assert abs(E - (mf.epskin + mf.epsH + mf.epsF)) < 1e-12


@pytest.mark.parametrize("which", ["hartree", "fock"])
def test_scf_single_channel_vs_oracle(pkg, gpu, mfm, hf, which):
    h, v, sub = pkg.models.honeycomb_hubbard(2.5)
    rho0 = mfm.staggered_guess(v, sub, m=0.3, noise=1e-3, seed=3)
    ks = pkg.regulargrid(nk=15 ** 2)
    fn = mfm.solvehartree if which == "hartree" else mfm.solvefock
    out = fn(h, v, rho0, 0.5, ks, T=0.02, beta=0.5, tol=1e-6, iterations=400, executor=gpu)
    ref = hf.scf(h.data, v.data, rho0.data, 0.5, ks, T=0.02, beta=0.5, tol=1e-6, iterations=400,
                 hartree=which == "hartree", fock=which == "fock")
    _compare_scf(pkg, mfm, out, ref, sub, T=0.02, ks=ks)


def test_scf_neel_order_graphene_U4(pkg, gpu, mfm, lat):
    L = lat.honeycomb()
    h = lat.graphene(L, mode="spinhalf")
    v = lat.gethubbard(L, mode="σx", a=0.5, U=4.0)
    sub = L.label("sublattice")
    rho0 = mfm.initialguess(v, "antiferro", sublattice=sub)
    rho, E, mf, conv, res = mfm.solvehartreefock(h, v, rho0, 0.5, klin=30, iterations=800, tol=1e-5,
                                                 T=0.01, beta=0.2, executor=gpu)
    assert conv
    a, b = _sub_mags(mfm, rho, sub)
    assert abs(np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b) + 1) < 1e-8


def test_scf_small_U_limit(pkg, gpu, mfm, lat):
    L = lat.honeycomb()
    h = pkg.models.addspin(lat.nearestneighbor(L, -1.0), "spinhalf")
    v = lat.gethubbard(L, mode="σx", a=0.5, U=1e-6)
    rho0 = mfm.initialguess(v, "antiferro", sublattice=L.label("sublattice"))
    rho, E, mf, conv, res = mfm.solvehartreefock(h, v, rho0, 0.5, klin=30, iterations=800, tol=1e-5,
                                                 T=0.0, beta=0.7, executor=gpu)
    assert conv
    ks = np.random.default_rng(3).random((2, 100))
    b, _ = pkg.api.bandmatrix(mf.hMF, ks)
    b = b - mf.mu
    rs = [(1, 0), (-1, 0), (0, 1), (0, -1), (1, -1), (-1, 1)]
    up = np.sqrt(3 + sum(np.cos(2 * np.pi * (ks[0] * r[0] + ks[1] * r[1])) for r in rs))
    for n, sgn in enumerate([-1, -1, 1, 1]):
        assert np.mean(np.abs(b[n] - sgn * up)) < 1e-5


def test_scf_mixed_precision_matches_fp64(pkg, gpu, mfm):
    h, v, sub = pkg.models.honeycomb_hubbard(3.0)
    rho0 = mfm.staggered_guess(v, sub, m=0.1, noise=1e-3, seed=1)
    ks = pkg.regulargrid(nk=30 ** 2)
    kw = dict(T=0.01, beta=0.5, tol=1e-8, iterations=500)
    a = mfm.solvehartreefock(h, v, rho0, 0.5, ks, executor=pkg.api.to_executor("gpu"), **kw)
    b = mfm.solvehartreefock(h, v, rho0, 0.5, ks, executor=pkg.api.to_executor("gpu", precision="mixed"), **kw)
    assert a[3] and b[3]
    assert abs(a[1] - b[1]) < 1e-8
    for R in a[0].keys():
        assert np.max(np.abs(a[0][R] - b[0][R])) < 1e-8


def test_scf_errors(pkg, gpu, mfm):
    h, v, sub = pkg.models.honeycomb_hubbard(3.0)
    bad = pkg.Hops({(1, 0): np.zeros((4, 4), complex), (-1, 0): np.zeros((4, 4), complex)})
    with pytest.raises(ValueError):
        mfm.HartreeFock(h, v, executor=gpu)(bad)  # no zero key for the Hartree term
    nonherm = pkg.Hops({(0, 0): np.diag([1j, 1.0, 0.0, 0.0])})  # n^T V0 n = 6i
    with pytest.raises(ValueError):
        mfm.HartreeFock(h, v, executor=gpu)(nonherm)  # complex-valued energies
    rho0 = mfm.staggered_guess(v, sub, m=0.1, noise=1e-3, seed=1)
    with pytest.raises(RuntimeError):
        mfm.solvehartreefock(h, v, rho0, 0.5, klin=12, T=0.01, beta=0.1, tol=1e-12, iterations=2,
                             convergenceerror=True, executor=gpu)
    with pytest.warns(UserWarning):
        out = mfm.solvehartreefock(h, v, rho0, 0.5, klin=12, T=0.01, beta=0.1, tol=1e-12, iterations=2,
                                   executor=gpu)
    assert out[2].history.shape == (2, 3) and not out[3]


def _spinful_random(pkg, n_sites, seed, U):
    """n_sites x 2 spins: a random spinless model (x) sigma_0 plus onsite Hubbard (sigma_x form)."""
    h = pkg.models.addspin(pkg.models.random_tb(n_sites, rmax=1, seed=seed), "spinhalf")
    sx = np.array([[0, 1], [1, 0]], dtype=complex)
    v = pkg.Hops({(0, 0): U * np.kron(np.eye(n_sites), sx)})
    return h, v


def test_scf_batch_uneven_models(pkg, gpu, mfm, hf):
    # uneven models (N = 4 .. 40) in one batched call on several streams: every result equals
    # the single-model SCF bit for bit (same kernels, same order) and the oracle to 1e-8
    rng = np.random.default_rng(44)
    probs = []
    for i, n in enumerate([2, 20, 6, 12, 3, 9]):
        h, v = _spinful_random(pkg, n, 60 + i, U=float(rng.uniform(1, 4)))
        sub = np.arange(n) % 2
        rho0 = mfm.staggered_guess(v, sub, m=0.2, noise=1e-3, seed=i)
        probs.append((h, v, rho0, 0.5, pkg.regulargrid(nk=8 ** 2)))
    kw = dict(T=0.02, beta=0.5, tol=1e-9, iterations=300)
    out = mfm.solvehartreefock_batch(probs, nstreams=4, executor=gpu, **kw)
    for (h, v, rho0, fill, ks), (rho, E, info, conv, res) in zip(probs, out):
        single = mfm.solvehartreefock(h, v, rho0, fill, ks, executor=gpu, **kw)
        assert E == single[1] and conv == single[3] and info["iterations"] == single[2].history.shape[0]
        for R in rho.keys():
            assert np.array_equal(rho[R], single[0][R])
    h, v, rho0, fill, ks = probs[1]
    ref = hf.scf(h.data, v.data, rho0.data, fill, ks, **kw)
    assert abs(out[1][1] - ref["E"]) < 1e-8
    for R in ref["rho"]:
        assert np.max(np.abs(out[1][0][R] - ref["rho"][R])) < 1e-8


def test_scf_large_n_path_vs_oracle(pkg, gpu, mfm, hf):
    # 130 orbitals (65 sites x spin): above the large-N threshold, so every k pass of the SCF
    # (mu from the eigenvalue pass, rho) runs the blocked eigensolver; oracle to 1e-8
    h, v = _spinful_random(pkg, 65, 90, U=2.5)
    rho0 = mfm.staggered_guess(v, np.arange(65) % 2, m=0.2, noise=1e-3, seed=3)
    ks = pkg.regulargrid(nk=3 ** 2)
    kw = dict(T=0.05, beta=0.5, tol=1e-9, iterations=200)
    rho, E, info, conv, res = mfm.solvehartreefock(h, v, rho0, 0.5, ks, executor=gpu, **kw)
    ref = hf.scf(h.data, v.data, rho0.data, 0.5, ks, **kw)
    assert conv and ref["converged"]
    assert abs(E - ref["E"]) < 1e-8
    for R in ref["rho"]:
        assert np.max(np.abs(rho[R] - ref["rho"][R])) < 1e-8


@pytest.mark.parametrize("case", ["honeycomb_60x60", "spinful_random_N24"])
def test_scf_graph_replay_equals_host_loop(pkg, gpu, mfm, case):
    # the graph-driven loop (first iteration eager, then replays of an 8-iteration graph, with
    # host synchronisations between replays) against the host-driven loop: bitwise equal
    # density matrix, energies, chemical potential, residual and history, and the same count
    if case == "honeycomb_60x60":
        h, v, sub = pkg.models.honeycomb_hubbard(3.0)
        rho0 = mfm.staggered_guess(v, sub, m=0.1, noise=1e-3, seed=1)
        ks = pkg.regulargrid(nk=60 ** 2)
        kw = dict(T=0.01, beta=0.5, tol=1e-8, iterations=500)
    else:
        h, v = _spinful_random(pkg, 12, 77, U=2.2)
        rho0 = mfm.staggered_guess(v, np.arange(12) % 2, m=0.2, noise=1e-3, seed=4)
        ks = pkg.regulargrid(nk=12 ** 2)
        kw = dict(T=0.03, beta=0.4, tol=1e-10, iterations=300)
    out = {}
    try:
        for graphs in (True, False, True):
            gpu.set_graphs(graphs)
            out.setdefault(graphs, []).append(mfm.solvehartreefock(h, v, rho0, 0.5, ks, executor=gpu, **kw))
    finally:
        gpu.set_graphs(True)
    ref = out[False][0]
    assert ref[3] and ref[2].history.shape[0] > 16  # several graph replays
    for res in out[True]:
        assert res[1] == ref[1] and res[3] == ref[3] and res[4] == ref[4]
        assert res[2].mu == ref[2].mu and res[2].epsband == ref[2].epsband
        assert np.array_equal(res[2].history, ref[2].history)
        for R in ref[0].keys():
            assert np.array_equal(res[0][R], ref[0][R])


def test_scf_graph_iteration_limit(pkg, gpu, mfm):
    # a limit that falls inside a graph replay stops at exactly that many iterations
    h, v, sub = pkg.models.honeycomb_hubbard(3.0)
    rho0 = mfm.staggered_guess(v, sub, m=0.1, noise=1e-3, seed=1)
    ks = pkg.regulargrid(nk=24 ** 2)
    res = {}
    try:
        for graphs in (True, False):
            gpu.set_graphs(graphs)
            with pytest.warns(UserWarning):
                res[graphs] = mfm.solvehartreefock(h, v, rho0, 0.5, ks, T=0.01, beta=0.5, tol=1e-14, iterations=13,
                                                   executor=gpu)
    finally:
        gpu.set_graphs(True)
    assert res[True][2].history.shape == (13, 3) and not res[True][3]
    assert np.array_equal(res[True][2].history, res[False][2].history)
    for R in res[False][0].keys():
        assert np.array_equal(res[True][0][R], res[False][0][R])
