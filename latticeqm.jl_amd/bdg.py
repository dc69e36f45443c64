"""Bogoliubov-de Gennes (Nambu) operators on the GPU band path.

Mirrors LatticeQM.jl's Superconductivity module for the band path:

=========================  ===========================================================
this module                reference
=========================  ===========================================================
``BdGOperator``            BdGOperator(h[, Delta]) (Superconductivity/types.jl:43-71)
``addpairing``             addpairing! (types.jl:93-118)
``addelectronsector``      addelectronsector! (types.jl:120-130)
``addchemicalpotential``   Operators.addchemicalpotential!(::BdGOperator, mu)
                           (Superconductivity/operators.jl:14-17)
``electron``               electron(H) (operators.jl:40-47)
``getelectronsector``      Utils.getelectronsector / copyelectronsector (types.jl:160-170)
``getpairingsector``       getpairingsector (types.jl:172-181)
``getdensitymatrix``       Operators.getdensitymatrix!(::BdGOperator, ...) (green.jl:28-54)
``filling``                Spectrum.filling(::BdGOperator, ks, mu) (operators.jl:107-117)
``chemicalpotential``      Spectrum.chemicalpotential(::BdGOperator, ks, target)
                           (operators.jl:122-162)
=========================  ===========================================================

The Nambu Hamiltonian H_BdG(k) = [[h(k), Delta(k)], [Delta^H(k), -h(-k)^T]] is itself a Hops
of dimension 2N (block R: [[h_R, Delta_R], [Delta_{-R}^H, -conj(h_R)]]), so every density-matrix
and band pass runs the ordinary kernels at 2N; only the chemical potential, which enters as
-mu tau_z on the zero key, is applied on the device by the lqm_bdg_* entry points.
"""
from __future__ import annotations

import ctypes as _c

import numpy as np

from . import _lib
from .hops import Hops, regulargrid


def _blk(a, b, c, d):
    return np.block([[a, b], [c, d]])


class BdGOperator:
    """Nambu operator of electron dimension ``hopdim`` = N; ``h`` holds the 2N x 2N blocks."""

    def __init__(self, h0: Hops | None = None, delta: Hops | None = None, *, _nambu: Hops | None = None):
        if _nambu is not None:
            self.h = _nambu
            return
        if not h0.ishermitian():
            raise AssertionError("BdGOperator(h): the normal-state hopping `h` must be Hermitian "
                                 "(h(R)^H = h(-R)); hermitianize it first")
        z = np.zeros((h0.N, h0.N), dtype=np.complex128)
        self.h = Hops({R: _blk(t, z, z, -np.conj(t)) for R, t in h0.items()})
        if delta is not None:
            addpairing(self, delta)

    @classmethod
    def from_nambu(cls, h2: Hops) -> "BdGOperator":
        """Wrap 2N x 2N Nambu blocks as they are (e.g. a container for a density matrix on
        chosen keys; BdGOperator{T}(h) of the reference)."""
        if h2.N % 2:
            raise ValueError("Nambu blocks have even dimension")
        return cls(_nambu=h2)

    # the Hops interface of the reference's wrapper (types.jl:183-206)
    def __call__(self, k):
        return self.h(k)

    def keys(self):
        return self.h.keys()

    def __getitem__(self, R):
        return self.h[R]

    def __setitem__(self, R, v):
        self.h[R] = v

    def __contains__(self, R):
        return R in self.h

    @property
    def hopdim(self) -> int:
        """Electron-sector size N (TightBinding.hopdim(::BdGOperator), types.jl:193)."""
        return self.h.N // 2

    @property
    def d(self) -> int:
        return self.h.d

    def zerokey(self):
        return self.h.zerokey()

    def ishermitian(self) -> bool:
        return self.h.ishermitian()

    def copy(self) -> "BdGOperator":
        return BdGOperator(_nambu=self.h.copy())

    def zero(self) -> "BdGOperator":
        return BdGOperator(_nambu=self.h.zero())


def _addhops(H: BdGOperator, extra: Hops) -> None:
    for R, t in extra.items():
        if R in H.h:
            H.h.data[R] = H.h.data[R] + t
        else:
            H.h[R] = t


def addpairing(H: BdGOperator, delta: Hops) -> BdGOperator:
    """Add the pairing blocks [[0, Delta_R], [Delta_{-R}^H, 0]] at every key R of ``delta``;
    every R needs its -R partner (the values of the two are independent)."""
    if delta.N != H.hopdim:
        raise AssertionError("Mismatch between matrix sizes of BdG Operator and pairing matrices.")
    for R in delta.keys():
        mR = tuple(-x for x in R)
        if mR not in delta:
            raise RuntimeError(f"addpairing: pairing dict has key {R} but no partner {mR}")
    z = np.zeros((H.hopdim, H.hopdim), dtype=np.complex128)
    _addhops(H, Hops({R: _blk(z, t, delta[tuple(-x for x in R)].conj().T, z) for R, t in delta.items()}))
    if not H.h.ishermitian():
        raise AssertionError("addpairing: resulting BdG operator is not Hermitian")
    return H


def addelectronsector(H: BdGOperator, h: Hops) -> BdGOperator:
    """Add h to the electron block and -conj(h) to the hole block."""
    if h.N != H.hopdim:
        raise AssertionError("Mismatch between matrix sizes of BdG Operator and pairing matrices.")
    if not h.ishermitian():
        raise AssertionError("addelectronsector: the electron-sector hopping `h` must be Hermitian.")
    z = np.zeros((h.N, h.N), dtype=np.complex128)
    _addhops(H, Hops({R: _blk(t, z, z, -np.conj(t)) for R, t in h.items()}))
    return H


def addchemicalpotential(H: BdGOperator, mu: float) -> BdGOperator:
    """+mu on the electron diagonal and -mu on the hole diagonal of the zero key."""
    d = H.hopdim
    blk = np.diag(np.concatenate([np.full(d, mu), np.full(d, -mu)])).astype(np.complex128)
    _addhops(H, Hops({H.zerokey(): blk}))
    return H


def electron(H: BdGOperator) -> Hops:
    """Projector onto the electron sector of Nambu space."""
    d = H.hopdim
    P = np.zeros((2 * d, 2 * d), dtype=np.complex128)
    P[np.arange(d), np.arange(d)] = 1.0
    return Hops({H.zerokey(): P})


def getelectronsector(H: BdGOperator) -> Hops:
    d = H.hopdim
    return Hops({R: t[:d, :d] for R, t in H.h.items()})


def getpairingsector(H: BdGOperator) -> Hops:
    d = H.hopdim
    return Hops({R: t[:d, d:] for R, t in H.h.items()})


def _nambu_model(H: BdGOperator, executor):
    """The 2N-dimensional Nambu model on the device, with the zero key present."""
    from .api import GPUModel, to_executor

    if not isinstance(H, BdGOperator):
        raise TypeError("expected a BdGOperator")
    h = H.h
    if h.zerokey() not in h:
        h = h.copy()
        h[h.zerokey()] = np.zeros((h.N, h.N), dtype=np.complex128)
    return GPUModel(h, to_executor(executor or "gpu"))


def getdensitymatrix(rho: BdGOperator, H: BdGOperator, ks, mu: float = 0.0, T: float = 0.01, kweights=None,
                     executor=None) -> float:
    """Overwrite the Nambu blocks of ``rho`` (its keys) with the density matrix of H - mu tau_z and
    return the physical <H - mu N> = E_band / 2 + (Re tr h_e(0) - N mu) / 2."""
    from .api import _kpts, _weights, _vp

    m = _nambu_model(H, executor)
    k = _kpts(ks, m.d)
    w = _weights(ks, kweights, k.shape[1])
    keys = list(rho.h.keys())
    for key in keys:
        if rho.h[key].shape != (m.N, m.N):
            raise ValueError("rho blocks must be 2N x 2N Nambu blocks")
    nRr = len(keys)
    Rr = np.asfortranarray(np.array(keys, dtype=np.int32).T.reshape(m.d, nRr)) if nRr else np.zeros((m.d, 1), np.int32)
    out = np.zeros((m.N, m.N, max(nRr, 1)), dtype=np.complex128, order="F")
    ek = _c.c_double()
    _lib.call("lqm_bdg_densitymatrix", m.handle, k.shape[1], _vp(k), _vp(w), float(mu), float(T), nRr, _vp(Rr),
              _vp(out), _c.byref(ek))
    for r, key in enumerate(keys):
        rho.h.data[key] = np.array(out[:, :, r])
    return ek.value


def filling(H: BdGOperator, ks=None, mu: float = 0.0, T: float = 0.01, nk: int | None = None, kweights=None,
            executor=None) -> float:
    """Electron filling Re tr(rho_el(0)) / N of the full Nambu density matrix at mu;
    ``nk`` instead of ``ks`` uses regulargrid(nk=nk^2) (Spectrum.filling(H, mu; nk))."""
    from .api import _kpts, _weights, _vp

    if ks is None:
        ks = regulargrid(nk=(10 if nk is None else nk) ** 2)
    m = _nambu_model(H, executor)
    k = _kpts(ks, m.d)
    w = _weights(ks, kweights, k.shape[1])
    out = _c.c_double()
    _lib.call("lqm_bdg_filling", m.handle, k.shape[1], _vp(k), _vp(w), float(mu), float(T), _c.byref(out))
    return out.value


def chemicalpotential(H: BdGOperator, ks, target: float, T: float = 0.01, mu_lo: float = -100.0,
                      mu_hi: float = 100.0, kweights=None, executor=None, return_passes: bool = False):
    """mu with filling(H, ks, mu) = target, the filling taken from the full Nambu density
    matrix; bracket from the electron-sector bands (extrema -+ 1, clipped to [mu_lo, mu_hi])."""
    from .api import _kpts, _weights, _vp, bandmatrix

    m = _nambu_model(H, executor)
    k = _kpts(ks, m.d)
    w = _weights(ks, kweights, k.shape[1])
    bands, _ = bandmatrix(getelectronsector(H), k, executor=m.exec)
    lo = max(mu_lo, float(np.min(bands)) - 1.0)
    hi = min(mu_hi, float(np.max(bands)) + 1.0)
    mu = _c.c_double()
    passes = _c.c_int()
    _lib.call("lqm_bdg_chemicalpotential", m.handle, k.shape[1], _vp(k), _vp(w), float(target), float(T), lo, hi,
              _c.byref(mu), _c.byref(passes))
    return (mu.value, passes.value) if return_passes else mu.value
