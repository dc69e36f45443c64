"""GPU density matrix, DOS and chemical potential vs the FP64 oracle.

Tolerances: rho to 1e-9 absolute (entries are O(1)), E_band to 1e-10 relative,
DOS to 1e-10 relative, mu: T=0 exact (order statistic of the same eigenvalues to
1e-12), T>0 the defining equation getfilling(mu) = filling to 1e-12.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _zero_like(pkg, h, keys=None):
    keys = list(h.keys()) if keys is None else keys
    return pkg.Hops({R: np.zeros((h.N, h.N), complex) for R in keys})


def _rho_vs_oracle(pkg, oracle, h, ks, mu, T, kweights=None, keys=None, tol=1e-9):
    rho = _zero_like(pkg, h, keys)
    eb = pkg.api.getdensitymatrix(rho, h, ks, mu, T=T, kweights=kweights)
    R, Tt = h.pack()
    Rr = np.array(list(rho.keys()), dtype=np.int32).T
    ref, eb_ref = oracle.densitymatrix(R, Tt, ks, Rr, mu=mu, Temp=T, kweights=kweights)
    err = max(np.max(np.abs(rho[tuple(Rr[:, r])] - ref[:, :, r])) for r in range(Rr.shape[1]))
    assert err < tol, err
    assert abs(eb - eb_ref) <= 1e-10 * max(1.0, abs(eb_ref))
    return rho, eb


@pytest.fixture(params=[(0, 0, -1), (1, 1, -1), (0, 1, -1), (1, 0, -1), (2, 0, -1), (2, 1, -1), (3, 0, -1),
                        (4, 1, -1), (0, 0, 0), (2, 0, 1)], ids=lambda v: f"tri{v[0]}-rho{v[1]}-bt{v[2]}")
def variants(request, gpu):
    """Force each (tridiagonalization, rho_k, back-transformation) kernel variant, then restore."""
    for kind, v in zip((0, 1, 4), request.param):
        gpu.set_variant(kind, v)
    yield request.param
    for kind in (0, 1, 4):
        gpu.set_variant(kind, -1)


@pytest.mark.parametrize("N,nk,mu,T", [(2, 36, 0.0, 0.01), (4, 64, 0.3, 0.0), (8, 49, -0.2, 0.05),
                                       (33, 25, 0.0, 0.02), (64, 64, 0.0, 0.02), (96, 9, 0.1, 0.02)])
def test_rho_random_models(pkg, oracle, gpu, variants, N, nk, mu, T):
    h = pkg.models.random_tb(N, rmax=1 if N > 8 else 2, seed=N)
    _rho_vs_oracle(pkg, oracle, h, pkg.regulargrid(nk=nk), mu, T)


@pytest.mark.parametrize("mu", [-1e3, 1e3, -1.5])
def test_rho_empty_full_and_partly_occupied_chunks(pkg, oracle, gpu, variants, mu):
    # occupations enter rho_k in chunks of 16 bands, and a chunk whose occupations are all
    # below 2^-70 is dropped: mu below every band (rho = 0), above every band (rho_k = I)
    # and a mu that leaves whole chunks empty beside partly filled ones
    h = pkg.models.random_tb(48, rmax=1, seed=61)
    _rho_vs_oracle(pkg, oracle, h, pkg.regulargrid(nk=25), mu, 0.02)


def test_rho_on_a_callers_stream(pkg, gpu):
    # the context on a caller's stream, with its internal side stream (the compact-WY T
    # factors, taken for chunks of >= 2048 k-points) forked from and joined back to it: the
    # same bits as on the context's own stream (every reduction has a fixed order)
    import torch

    h = pkg.models.random_tb(64, rmax=1, seed=71)
    ks = pkg.regulargrid(nk=48 * 48)

    def run():
        rho = _zero_like(pkg, h)
        eb = pkg.api.getdensitymatrix(rho, h, ks, 0.0, T=0.02, executor=gpu)
        return rho, eb

    ref_rho, ref_eb = run()
    s = torch.cuda.Stream()
    gpu.set_stream(s.cuda_stream)
    try:
        with torch.cuda.stream(s):
            rho, eb = run()
        s.synchronize()
    finally:
        gpu.set_stream(None)
    assert eb == ref_eb
    for key in ref_rho.keys():
        assert np.array_equal(rho[key], ref_rho[key])


def test_rho_degenerate_spinful_all_variants(pkg, oracle, gpu, variants):
    h = pkg.models.addspin(pkg.models.random_tb(5, rmax=1, seed=3), "spinhalf")
    _rho_vs_oracle(pkg, oracle, h, pkg.regulargrid(nk=36), 0.1, 0.0)


def test_autotune_picks_valid_variants(pkg, oracle, gpu):
    import torch

    h = pkg.models.random_tb(16, rmax=1, seed=4)
    ks = pkg.regulargrid(nk=256)
    m = pkg.api.GPUModel(h, gpu)
    ks_d = torch.from_numpy(np.ascontiguousarray(ks.T)).cuda()
    chosen = pkg.api.autotune(m, ks_d.data_ptr(), ks.shape[1])
    assert chosen[0] in (0, 1, 2, 3, 4) and chosen[1] in (0, 1) and chosen[3] in (0, 1, 2, 3)
    _rho_vs_oracle(pkg, oracle, h, ks, 0.0, 0.05)


def test_autotune_cache_roundtrip(pkg, gpu, tmp_path):
    import torch

    h = pkg.models.random_tb(24, rmax=1, seed=8)
    ks = pkg.regulargrid(nk=256)
    ex = pkg.api.GPUExec(0)  # contexts of their own (to_executor shares one per device)
    m = pkg.api.GPUModel(h, ex)
    ks_d = torch.from_numpy(np.ascontiguousarray(ks.T)).cuda()
    chosen = pkg.api.autotune(m, ks_d.data_ptr(), ks.shape[1])
    path = tmp_path / "tuning.txt"
    ex.save_tuning(path)
    fresh = pkg.api.GPUExec(0)
    assert all(fresh.get_variant(kind, 24) == -1 for kind in chosen)
    fresh.load_tuning(path)
    assert {kind: fresh.get_variant(kind, 24) for kind in chosen} == chosen
    # another architecture's file, and an out-of-range entry, are refused and change nothing
    lines = path.read_text().splitlines()
    other = tmp_path / "other.txt"
    other.write_text("\n".join(["lqm-tuning 1 gfx000"] + lines[1:]) + "\n")
    bad = tmp_path / "bad.txt"
    bad.write_text(lines[0] + "\n0 24 9\n")
    third = pkg.api.GPUExec(0)
    for f in (other, bad, tmp_path / "missing.txt"):
        with pytest.raises(Exception):
            third.load_tuning(f)
    assert all(third.get_variant(kind, 24) == -1 for kind in chosen)


def test_rho_zeeman_graphene_T0(pkg, oracle, gpu):
    # the test_operators.jl:91-127 configuration
    h = pkg.models.addspin(pkg.models.honeycomb_nn(-1.0), "spinhalf")
    sy = np.array([[0, -1j], [1j, 0]])
    h[(0, 0)] = h[(0, 0)] + np.kron(np.eye(2), 0.4 * sy)
    nk = 9
    ks = np.array([[i / nk, j / nk] for i in range(nk) for j in range(nk)]).T
    _rho_vs_oracle(pkg, oracle, h, ks, -0.5, 0.0)


def test_rho_weights_and_extra_keys(pkg, oracle, gpu):
    h = pkg.models.random_tb(6, rmax=1, seed=6)
    ks = pkg.regulargrid(nk=16)
    w = np.random.default_rng(0).random(16)
    keys = [(0, 0), (2, -1), (3, 3), (-1, 0)]
    _rho_vs_oracle(pkg, oracle, h, ks, 0.0, 0.03, kweights=w, keys=keys)


def test_rho_many_keys(pkg, oracle, gpu):
    # more keys than one register group of the phase contraction
    h = pkg.models.random_tb(5, rmax=3, seed=8)  # 49 keys
    _rho_vs_oracle(pkg, oracle, h, pkg.regulargrid(nk=16), 0.1, 0.05)


def test_rho_invariants_cfg2_shape(pkg, gpu):
    # N=64, 25 keys, at a grid the oracle check above does not cover: size-free identities
    h = pkg.models.random_tb(64, rmax=2, seed=2)
    ks = pkg.regulargrid(nk=32 ** 2)
    rho = _zero_like(pkg, h)
    eb = pkg.api.getdensitymatrix(rho, h, ks, 0.0, T=0.02)
    for R in rho.keys():
        mR = tuple(-x for x in R)
        assert np.max(np.abs(rho[mR] - rho[R].conj().T)) < 1e-12
    pair = sum(np.trace(rho[R] @ h[tuple(-x for x in R)]) for R in rho.keys())
    # keys of rho = keys of h cover only part of the Fourier content, so E_band is
    # sum over the complete key set; with all |R| <= 2 blocks of h present the pairing is exact
    assert abs(np.real(pair) - eb) < 1e-10 * max(1, abs(eb))


def test_dos_vs_oracle(pkg, oracle, gpu):
    h = pkg.models.honeycomb_nn(-1.0)
    R, T = h.pack()
    ks = pkg.regulargrid(nk=60 ** 2)
    om = np.linspace(-3.05, 3.05, 401)
    dos = pkg.api.getdos(h, om, ks, Gamma=0.05)
    ref = oracle.dos(R, T, ks, om, 0.05)
    assert np.max(np.abs(dos - ref)) < 1e-10 * np.max(ref)
    om2, dos2 = pkg.api.getdos(h, -3.05, 3.05, 401, klin=60, Gamma=0.05)
    assert np.allclose(om2, om) and np.max(np.abs(dos2 - ref)) < 1e-10 * np.max(ref)
    w = np.full(ks.shape[1], 1.0 / ks.shape[1])
    dw = pkg.api.getdos(h, om, ks, Gamma=0.05, kweights=w)
    assert np.max(np.abs(dw - ref)) < 1e-10 * np.max(ref)


@pytest.mark.parametrize("filling,T", [(0.5, 0.0), (0.25, 0.0), (1.0, 0.0), (0.0, 0.0), (0.5, 0.01), (0.37, 0.2),
                                       (0.02, 0.05), (0.98, 0.05), (0.5, 0.001)])
def test_chemicalpotential_vs_oracle(pkg, oracle, gpu, filling, T):
    h = pkg.models.random_tb(6, rmax=1, seed=12)
    ks = pkg.regulargrid(nk=49)
    R, Tt = h.pack()
    bands = oracle.bands(R, Tt, ks)
    mu = pkg.api.chemicalpotential(h, ks, filling, T=T)
    if T == 0:
        assert abs(mu - oracle.chemicalpotential(bands, filling, 0.0)) < 1e-12
    else:
        assert abs(oracle.getfilling(bands, mu, T) - filling) < 1e-12
        assert abs(mu - oracle.chemicalpotential(bands, filling, T)) < 1e-10


def test_chemicalpotential_in_a_gap(pkg, oracle, gpu):
    # half filling inside a gap of width > 2 at kT = 0.01: the filling is flat to rounding
    # across the gap, so mu must lie in the gap and solve the defining equation
    base = pkg.models.random_tb(4, rmax=1, seed=3)
    h = pkg.Hops({key: 0.1 * base[key] for key in base.keys()})
    h[(0, 0)] = h[(0, 0)] + np.diag([2.0, 2.0, -2.0, -2.0])
    ks = pkg.regulargrid(nk=36)
    R, Tt = h.pack()
    bands = oracle.bands(R, Tt, ks)
    vals = np.sort(bands.ravel())
    lo, hi = vals[vals.size // 2 - 1], vals[vals.size // 2]
    assert hi - lo > 2.0
    mu = pkg.api.chemicalpotential(h, ks, 0.5, T=0.01)
    assert lo < mu < hi
    assert abs(oracle.getfilling(bands, mu, 0.01) - 0.5) < 1e-12


def test_setfilling_graphene(pkg, gpu):
    # test_spectrum_graphene.jl: half filling of NN graphene puts mu at the Dirac point
    h = pkg.models.honeycomb_nn(-1.0)
    mu = pkg.api.setfilling(h, 0.5, nk=100)
    assert abs(mu) < 1e-12


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_small_models_one_thread_per_k(pkg, oracle, N):
    # N <= 4 runs on one thread per k-point (Jacobi in registers); the same passes with a forced
    # variant take the per-k workgroup kernels: both against the oracle
    if N == 4:
        h = pkg.models.addspin(pkg.models.random_tb(2, rmax=1, seed=31), "spinhalf")  # exact pairs
    else:
        h = pkg.models.random_tb(N, rmax=1, seed=30 + N)
    ks = pkg.regulargrid(nk=100)
    R, T = h.pack()
    ex = pkg.api.GPUExec(0)
    forced = pkg.api.GPUExec(0)
    forced.set_variant(0, 0)
    bref = oracle.bands(R, T, ks)
    for e in (ex, forced):
        b, _ = pkg.api.bandmatrix(h, ks, executor=e)
        assert np.max(np.abs(b - bref)) < 1e-12 * max(1.0, np.max(np.abs(bref)))
        _, V = pkg.api.eigen(h, ks[:, :7], executor=e)
        for j in range(7):
            Hk = h(ks[:, j])
            Hk = np.triu(Hk) + np.triu(Hk, 1).conj().T
            assert np.max(np.abs(Hk @ V[:, :, j] - V[:, :, j] * b[:, j])) < 1e-12 * max(1.0, np.max(np.abs(bref)))
            assert np.max(np.abs(V[:, :, j].conj().T @ V[:, :, j] - np.eye(h.N))) < 1e-13
        w = np.linspace(0.5, 1.5, ks.shape[1]) / ks.shape[1]
        for mu, Tk, kw in ((0.1, 0.05, None), (0.0, 0.0, None), (-0.2, 0.02, w)):
            rho = pkg.Hops({key: np.zeros((h.N, h.N), complex) for key in h.keys()})
            eb = pkg.api.getdensitymatrix(rho, h, ks, mu, T=Tk, kweights=kw, executor=e)
            ref, ebr = oracle.densitymatrix(R, T, ks, R, mu=mu, Temp=Tk, kweights=kw)
            for r, key in enumerate(h.keys()):
                assert np.max(np.abs(rho[key] - ref[:, :, r])) < 1e-12, (mu, Tk, key)
            assert abs(eb - ebr) < 1e-12 * max(1.0, abs(ebr))


def _rho_keys(pkg, h, ks, keys, mu, T, ex):
    # lqm_densitymatrix on an explicit key list (repeats allowed), through the C ABI
    import ctypes

    api = pkg.api
    m = api._model(h, ex)
    k = api._kpts(ks, m.d)
    w = api._weights(ks, None, k.shape[1])
    Rr = np.asfortranarray(np.array(keys, dtype=np.int32).T)
    rho = np.zeros((m.N, m.N, len(keys)), dtype=np.complex128, order="F")
    eb = ctypes.c_double()
    api._lib.call("lqm_densitymatrix", m.handle, k.shape[1], api._vp(k), api._vp(w), float(mu), float(T),
                  len(keys), api._vp(Rr), api._vp(rho), ctypes.byref(eb))
    return rho, eb.value


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("N", [3, 7, 33, 64, 96])
def test_rho_hermitian_contraction(pkg, oracle, prec, N):
    # a key list closed under R -> -R (no repeats) is contracted over the upper triangle with the
    # lower one taken from rho(-R); one key without its negation, or a repeated key, takes the
    # full contraction.  All three vs the oracle; the shared keys agree between the modes
    h = pkg.models.random_tb(N, rmax=1, seed=910 + N)
    ks = pkg.regulargrid(nk=16 if N < 90 else 9)
    ex = pkg.api.to_executor("gpu", precision=prec)
    closed = list(h.keys())
    lists = {"closed": closed, "extra": closed + [(2, -1)], "repeat": closed + [closed[1]]}
    R, Tt = h.pack()
    out = {}
    for name, keys in lists.items():
        rho, eb = _rho_keys(pkg, h, ks, keys, 0.05, 0.02, ex)
        ref, eb_ref = oracle.densitymatrix(R, Tt, ks, np.array(keys, dtype=np.int32).T, mu=0.05, Temp=0.02)
        assert np.max(np.abs(rho - ref)) < 1e-9, name
        assert abs(eb - eb_ref) <= 1e-10 * max(1.0, abs(eb_ref))
        out[name] = rho
    n = len(closed)
    scale = np.max(np.abs(out["closed"]))
    for name in ("extra", "repeat"):
        assert np.max(np.abs(out[name][:, :, :n] - out["closed"])) < 1e-13 * max(1.0, scale), name
    assert np.array_equal(out["repeat"][:, :, n], out["repeat"][:, :, 1])
    # the closed list's result: rho(-R) = rho(R)^H, exactly off the diagonal (one sum feeds both)
    off = ~np.eye(N, dtype=bool)
    for r, key in enumerate(closed):
        q = closed.index(tuple(-x for x in key))
        A, B = out["closed"][:, :, r], out["closed"][:, :, q].conj().T
        assert np.array_equal(A[off], B[off])
        assert np.max(np.abs(A - B)) < 1e-14 * max(1.0, scale)
