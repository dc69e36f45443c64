"""Hartree-Fock self-consistency on the GPU, mirroring LatticeQM's ``Meanfield`` module.

=========================  =====================================================
this module                reference
=========================  =====================================================
``HartreeFock``            Meanfield.HartreeFock (Meanfield/hartreefock.jl:143-221)
``solvehartreefock``       Meanfield.solvehartreefock (Meanfield/selfconsistent.jl:19)
``solvehartree/solvefock`` selfconsistent.jl:20-21
``sanitize``               Meanfield.sanitize! (selfconsistent.jl:106-113)
``initialguess``           Meanfield.initialguess (Meanfield/initialguess.jl:193-216)
``enrichkeys``             Meanfield.enrichkeys! (initialguess.jl:255-274)
``gethubbard``             Operators.gethubbard (Operators/interactions.jl:8-28)
``trace`` / ``expval``     Operators.trace / expval (Operators/trace.jl:16-47)
=========================  =====================================================

The whole self-consistent loop (H_MF[rho], mu, rho, energies, mixing, residual) runs in
liblqmgpu.so; this module packs the inputs and returns the reference's result tuple
``(rho, E_GS, mf, converged, residual)``.
"""
from __future__ import annotations

import ctypes
import warnings
from typing import Callable, Optional

import numpy as np

from . import _lib
from .api import _kpts, _vp, _weights, to_executor
from .hops import Hops, regulargrid

_c = ctypes


def _pack_keys(H: Hops, N: int, d: int):
    R, T = H.pack()
    if T.shape[0] != N:
        raise ValueError("all operators must have the same orbital count")
    if R.shape[0] != d:
        raise ValueError("all operators must have the same lattice dimension")
    return R, T


def sanitize(rho: Hops) -> Hops:
    """Hermitianize rho in place unless it already is Hermitian (selfconsistent.jl:106-113)."""
    if not rho.ishermitian():
        rho.hermitianize()
    return rho


class HartreeFock:
    """Mean-field functional h + Fock[rho] + Hartree[rho] and its scalar energies.

    Calling ``hf(rho)`` sets ``hMF`` (a Hops on keys(h), then keys(v) missing from h,
    then the zero key) and ``epsH`` / ``epsF``; the SCF also fills ``mu``, ``epsband`` and
    ``epskin``.  Physical-sign energies: eps_H = 1/2 n^T V0 n, eps_F = -1/2 sum v|rho|^2.
    """

    def __init__(self, h: Hops, v: Hops, mu: float = 0.0, hartree: bool = True, fock: bool = True,
                 executor=None):
        self.h, self.v, self.mu = h, v, float(mu)
        self.hartree, self.fock = bool(hartree), bool(fock)
        self.exec = to_executor(executor or "gpu")
        self.N, self.d = h.N, h.d
        self.V0 = sum(np.asarray(b) for b in v.values()) if len(v) else np.zeros((self.N, self.N), complex)
        self.hMF: Optional[Hops] = None
        self.epsH = self.epsF = self.epsband = self.epskin = 0.0

    @property
    def flags(self) -> int:
        return (_lib.LQM_HF_HARTREE if self.hartree else 0) | (_lib.LQM_HF_FOCK if self.fock else 0)

    def _hv(self):
        Rh, Th = _pack_keys(self.h, self.N, self.d)
        if len(self.v):
            Rv, Tv = _pack_keys(self.v, self.N, self.d)
        else:
            Rv = np.zeros((self.d, 1), np.int32)
            Tv = np.zeros((self.N, self.N, 1), np.complex128, order="F")
        return Rh, Th, Rv, Tv, len(self.v)

    def _outbufs(self):
        cap = len(self.h) + len(self.v) + 1
        nK = _c.c_int()
        RK = np.zeros((self.d, cap), dtype=np.int32, order="F")
        hmf = np.zeros((self.N, self.N, cap), dtype=np.complex128, order="F")
        return nK, RK, hmf

    def _set_hmf(self, nK, RK, hmf) -> None:
        self.hMF = Hops({tuple(int(x) for x in RK[:, q]): hmf[:, :, q] for q in range(nK.value)})

    def __call__(self, rho: Hops) -> "HartreeFock":
        Rh, Th, Rv, Tv, nRv = self._hv()
        Rr, Tr = _pack_keys(rho, self.N, self.d)
        nK, RK, hmf = self._outbufs()
        en = (_c.c_double * 2)()
        _lib.call("lqm_hartreefock_meanfield", self.exec.handle, self.N, self.d, Rh.shape[1], _vp(Rh),
                  _vp(Th), nRv, _vp(Rv), _vp(Tv), Rr.shape[1], _vp(Rr), _vp(Tr), self.flags,
                  _c.byref(nK), _vp(RK), _vp(hmf), en)
        self._set_hmf(nK, RK, hmf)
        self.epsH, self.epsF = en[0], en[1]
        return self

    def doublecounting(self) -> float:
        """eps_H + eps_F (hartreefock.jl:239-240)."""
        return self.epsH + self.epsF

    def grandpotential(self, rho: Hops, ks, mu: float = 0.0, T: float = 0.05, kweights=None,
                       gradient: bool = False):
        """Omega of H_MF = hf(rho) at (mu, T) on ks; with ``gradient=True`` also the
        derivatives at fixed rho with respect to the mean-field parameters, as Hops on the keys
        of h and of v: (Omega, dh, dv) with dh[R] = dOmega/dh_R (element-wise, = rho_MF(-R)^T)
        and dv[L] = dOmega/dv_L (Fock and Hartree chain rule through H_MF)."""
        Rh, Th, Rv, Tv, nRv = self._hv()
        Rr, Tr = _pack_keys(rho, self.N, self.d)
        k = _kpts(ks, self.d)
        w = _weights(ks, kweights, k.shape[1])
        om = _c.c_double()
        gh = np.zeros((self.N, self.N, Rh.shape[1]), dtype=np.complex128, order="F") if gradient else None
        gv = np.zeros((self.N, self.N, max(nRv, 1)), dtype=np.complex128, order="F") if gradient else None
        _lib.call("lqm_hartreefock_grandpotential", self.exec.handle, self.N, self.d, Rh.shape[1], _vp(Rh), _vp(Th),
                  nRv, _vp(Rv), _vp(Tv), Rr.shape[1], _vp(Rr), _vp(Tr), self.flags, k.shape[1], _vp(k), _vp(w),
                  float(mu), float(T), _c.byref(om), _vp(gh), _vp(gv))
        if not gradient:
            return om.value
        dh = Hops({key: np.array(gh[:, :, r]) for r, key in enumerate(self.h.keys())})
        dv = Hops({key: np.array(gv[:, :, r]) for r, key in enumerate(self.v.keys())})
        return om.value, dh, dv

    def grandpotential_rrule(self, rho: Hops, ks, mu: float = 0.0, T: float = 0.05, kweights=None):
        """(Omega, pullback): pullback(Omega_bar) returns the cotangents (of h, of v) in the
        convention of ChainRulesCore for complex inputs, Omega_bar * conj(dOmega/dx)."""
        om, dh, dv = self.grandpotential(rho, ks, mu, T, kweights, gradient=True)

        def pullback(dbar: float = 1.0):
            return (Hops({R: dbar * np.conj(b) for R, b in dh.items()}),
                    Hops({R: dbar * np.conj(b) for R, b in dv.items()}))

        return om, pullback


def solveselfconsistent(rho_init: Hops, mf: HartreeFock, filling: float, ks=None, *,
                        klin: int | None = None, iterations: int = 500, tol: float = 1e-7,
                        T: float = 0.0, beta: float = 1.0, convergenceerror: bool = False,
                        log_callback: Callable | None = None, kweights=None, **kw):
    """Copying wrapper: iterate ``mf`` from ``rho_init`` to self-consistency at ``filling``.

    ``ks`` (d x Nk points or a Mesh) or ``klin`` (regulargrid(nk=klin^d)) define the
    k-grid.  Linear mixing ``beta`` (``β`` accepted); stop at residual < ``tol``.
    Returns ``(rho, E_GS, mf, converged, residual)``; ``mf.history`` holds
    (E, residual, mu) per iteration, and ``log_callback(iter, E, residual)`` is called
    for each of them once the device loop has finished.
    """
    beta = float(kw.pop("β", beta))
    if kw:
        raise TypeError(f"unexpected keyword arguments {sorted(kw)}")
    if ks is None:
        if klin is None:
            raise ValueError("supply ks or klin")
        ks = regulargrid(nk=klin ** mf.d, dim=mf.d)
    rho0 = sanitize(rho_init.copy())
    k = _kpts(ks, mf.d)
    w = _weights(ks, kweights, k.shape[1])
    Rh, Th, Rv, Tv, nRv = mf._hv()
    Rr, Tr = _pack_keys(rho0, mf.N, mf.d)
    nK, RK, hmf = mf._outbufs()
    result = (_c.c_double * 8)()
    conv = _c.c_int()
    hist = np.zeros((3, max(int(iterations), 1)), order="F")
    _lib.call("lqm_hartreefock_scf", mf.exec.handle, mf.N, mf.d, Rh.shape[1], _vp(Rh), _vp(Th), nRv,
              _vp(Rv), _vp(Tv), Rr.shape[1], _vp(Rr), _vp(Tr), k.shape[1], _vp(k), _vp(w),
              float(filling), float(T), beta, float(tol), int(iterations), mf.flags, result,
              _c.byref(conv), _vp(hist), _c.byref(nK), _vp(RK), _vp(hmf))
    mf._set_hmf(nK, RK, hmf)
    E, mf.mu, mf.epsband, mf.epsH, mf.epsF, mf.epskin, residual, nit = list(result)
    nit = int(nit)
    mf.history = hist[:, :nit].T.copy()
    if log_callback is not None:
        for i in range(nit):
            log_callback(i + 1, mf.history[i, 0], mf.history[i, 1])
    converged = bool(conv.value)
    if not converged:
        if convergenceerror:
            raise RuntimeError("Convergence error.")
        warnings.warn(f"Did not converge to the requested tolerance. Residual: {residual}")
    rho = Hops({tuple(int(x) for x in Rr[:, r]): Tr[:, :, r] for r in range(Rr.shape[1])})
    return sanitize(rho), E, mf, converged, residual


def solvehartreefock(h: Hops, v: Hops, rho_init: Hops, filling: float, ks=None, executor=None, **kw):
    """solvehartreefock(h, v, rho_init, filling, [ks]; klin, iterations, tol, T, beta)."""
    return solveselfconsistent(rho_init, HartreeFock(h, v, executor=executor), filling, ks, **kw)


def solvehartree(h: Hops, v: Hops, rho_init: Hops, filling: float, ks=None, executor=None, **kw):
    return solveselfconsistent(rho_init, HartreeFock(h, v, fock=False, executor=executor), filling, ks, **kw)


def solvefock(h: Hops, v: Hops, rho_init: Hops, filling: float, ks=None, executor=None, **kw):
    return solveselfconsistent(rho_init, HartreeFock(h, v, hartree=False, executor=executor), filling, ks, **kw)


# ---- interactions, initial guesses, order parameters (host set-up and readout) -------
_SX = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_SY = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_SZ = np.array([[1, 0], [0, -1]], dtype=np.complex128)


def gethubbard(nsites: int, U: float = 1.0, mode: str = "sigmax", d: int = 2) -> Hops:
    """Onsite Hubbard interaction U * heaviside(a - r) with a below the nearest-neighbour
    distance (the gethubbard(lat; a=0.5) case of interactions.jl:8, :18-28): a single zero
    key, then addspin(mode): "sigmax" gives U * kron(I, sigma_x) (types.jl:258-270)."""
    base = U * np.eye(nsites, dtype=np.complex128)
    if mode in ("sigmax", "σx"):
        blk = np.kron(base, _SX)
    elif mode in ("spinhalf", "sigma0", "σ0"):
        blk = np.kron(base, np.eye(2))
    elif mode in ("nospin", "id"):
        blk = base
    else:
        raise ValueError(f"unknown spin mode {mode!r}")
    return Hops({(0,) * d: blk})


def spindensitymatrix(direction=(0.0, 0.0, 1.0)) -> np.ndarray:
    """1/2 (sigma_0 + d . sigma), d normalised when |d| > 1 (initialguess.jl:18-35)."""
    dvec = np.asarray(direction, dtype=np.float64)
    n = np.linalg.norm(dvec)
    if n > 1:
        dvec = dvec / n
    return 0.5 * (np.eye(2) + dvec[0] * _SX + dvec[1] * _SY + dvec[2] * _SZ)


def initialguess(v: Hops, mode: str = "antiferro", sublattice=None, direction="up") -> Hops:
    """Density matrix on keys(v) (initialguess.jl:193-216): every block zero except the
    zero key.  "ferro": I_n (x) rho_s; "antiferro": P_A (x) rho_s + P_B (x) (sigma_0 - rho_s)
    with ``sublattice`` the 0/1 label of each site (setantiferro!, :183-191)."""
    N = v.N
    if N % 2:
        raise ValueError("the hopping matrix must have even dimension (spinful)")
    n = N // 2
    rs = spindensitymatrix({"up": (0, 0, 1), "down": (0, 0, -1), "x": (1, 0, 0), "y": (0, 1, 0)}
                           .get(direction, direction) if isinstance(direction, str) else direction)
    rho = Hops({R: np.zeros((N, N), dtype=np.complex128) for R in v.keys()})
    zk = v.zerokey()
    if mode == "ferro":
        rho[zk] = np.kron(np.eye(n), rs)
    elif mode == "antiferro":
        if sublattice is None:
            raise ValueError("antiferro needs the sublattice label of each site")
        lab = np.asarray(sublattice)
        PA = np.diag((lab == 0).astype(float))
        PB = np.diag((lab == 1).astype(float))
        rho[zk] = np.kron(PA, rs) + np.kron(PB, np.eye(2) - rs)
    else:
        raise ValueError(f"unrecognized mode {mode!r}")
    return rho


def staggered_guess(v: Hops, sublattice, m: float = 0.1, noise: float = 1e-3, seed: int = 1) -> Hops:
    """Staggered-magnetisation start: P_A (x) 1/2(s0 + m sz) + P_B (x) 1/2(s0 - m sz) on the
    zero key plus Hermitian noise with entries U(-noise, noise) (real and imaginary parts,
    numpy default_rng(seed)); the other keys of v stay zero."""
    N = v.N
    lab = np.asarray(sublattice)
    PA = np.diag((lab == 0).astype(float))
    PB = np.diag((lab == 1).astype(float))
    blk = np.kron(PA, 0.5 * (np.eye(2) + m * _SZ)) + np.kron(PB, 0.5 * (np.eye(2) - m * _SZ))
    rng = np.random.default_rng(seed)
    X = rng.uniform(-noise, noise, (N, N)) + 1j * rng.uniform(-noise, noise, (N, N))
    blk = blk + (X + X.conj().T) / 2
    rho = Hops({R: np.zeros((N, N), dtype=np.complex128) for R in v.keys()})
    rho[v.zerokey()] = blk
    return rho


def enrichkeys(rho: Hops, klin: int) -> Hops:
    """Add zero blocks for every key of the box {-klin//2 .. klin//2}^d (initialguess.jl:255-274)."""
    if klin % 2 == 0:
        warnings.warn("enrichkeys: even klin leaves duplicate boundary keys; use odd klin")
    L = klin // 2
    import itertools

    for key in itertools.product(range(-L, L + 1), repeat=rho.d):
        if key not in rho:
            rho[key] = np.zeros((rho.N, rho.N), dtype=np.complex128)
    return rho


def trace(rho: Hops, O=None):
    """trace(rho) = tr rho[0]; trace(rho, O) = tr(rho[0] O) for a matrix O;
    trace(A, B) = sum_L tr(A_L B_{-L}) for two Hops (Operators/trace.jl:16-29)."""
    zk = rho.zerokey()
    if O is None:
        return np.trace(rho[zk])
    if isinstance(O, Hops):
        return sum(np.sum(O[tuple(-x for x in L)].T * A) for L, A in rho.items()
                   if tuple(-x for x in L) in O)
    return np.trace(rho[zk] @ np.asarray(O))


def expval(rho: Hops, O) -> complex:
    """<O> = sum_L tr(rho_L O_{-L}); a matrix O is an onsite operator (trace.jl:40-45)."""
    if not isinstance(O, Hops):
        O = Hops({rho.zerokey(): np.asarray(O, dtype=np.complex128)})
    return trace(rho, O)


def spin_operators(nsites: int):
    """[s_x, s_y, s_z] = I_n (x) sigma (spin the fast index; Operators/spin.jl:17-29)."""
    return [np.kron(np.eye(nsites), s) for s in (_SX, _SY, _SZ)]


def sublatticeprojection(sublattice, label: int, d: int = 2) -> np.ndarray:
    """kron(diag(sublattice == label), I_d) (Operators/sublattice.jl:3-21)."""
    lab = np.asarray(sublattice)
    return np.kron(np.diag((lab == label).astype(float)), np.eye(d))


def magnetization(rho: Hops, A=None) -> np.ndarray:
    """Real <s_x>, <s_y>, <s_z> (optionally projected by the onsite matrix A)."""
    n = rho.N // 2
    ops = spin_operators(n)
    if A is not None:
        ops = [A @ s for s in ops]
    return np.real(np.array([expval(rho, s) for s in ops]))


def solvehartreefock_batch(problems, T: float = 0.0, beta: float = 1.0, tol: float = 1e-7,
                           iterations: int = 500, hartree: bool = True, fock: bool = True,
                           nstreams: int = 4, executor=None, progress=None):
    """Many independent Hartree-Fock SCFs of uneven size in one call.

    ``problems``: sequence of (h, v, rho_init, filling, ks) tuples (ks: points, Mesh or an
    int klin).  They run concurrently on ``nstreams`` device streams, largest first.
    Returns, per problem, ``(rho, E_GS, info, converged, residual)`` with info holding mu,
    epsband, epsH, epsF, epskin and the iteration count.  ``progress(i, iterations,
    converged)``, if given, is called as problem i finishes (from the batch's threads).
    """
    ex = to_executor(executor or "gpu")
    keep = []
    arr = (_lib.HFProblem * len(problems))()
    d = None
    for i, (h, v, rho_init, filling, ks) in enumerate(problems):
        d = h.d if d is None else d
        if h.d != d:
            raise ValueError("all problems must have the same lattice dimension")
        if isinstance(ks, int):
            ks = regulargrid(nk=ks ** d, dim=d)
        N = h.N
        rho0 = sanitize(rho_init.copy())
        Rh, Th = _pack_keys(h, N, d)
        if len(v):
            Rv, Tv = _pack_keys(v, N, d)
        else:
            Rv, Tv = np.zeros((d, 1), np.int32), np.zeros((N, N, 1), np.complex128, order="F")
        Rr, Tr = _pack_keys(rho0, N, d)
        k = _kpts(ks, d)
        w = _weights(ks, None, k.shape[1])
        keep.append((Rh, Th, Rv, Tv, Rr, Tr, k, w))
        p = arr[i]
        p.N, p.nRh, p.Rh, p.Th = N, Rh.shape[1], Rh.ctypes.data, Th.ctypes.data
        p.nRv, p.Rv, p.Tv = len(v), Rv.ctypes.data, Tv.ctypes.data
        p.nRr, p.Rr, p.rho = Rr.shape[1], Rr.ctypes.data, Tr.ctypes.data
        p.Nk, p.ks, p.kweights = k.shape[1], k.ctypes.data, (w.ctypes.data if w is not None else None)
        p.filling = float(filling)
    flags = (_lib.LQM_HF_HARTREE if hartree else 0) | (_lib.LQM_HF_FOCK if fock else 0)
    cb = None
    if progress is not None:
        cb = _lib.PROGRESS_FN(lambda i, it, conv, _u: progress(int(i), int(it), bool(conv)))
        _lib.call("lqm_context_set_progress", ex.handle, ctypes.cast(cb, ctypes.c_void_p), None)
    try:
        _lib.call("lqm_hartreefock_scf_batch", ex.handle, int(d or 2), len(problems), arr, float(T), float(beta),
                  float(tol), int(iterations), flags, int(nstreams))
    finally:
        if cb is not None:
            _lib.call("lqm_context_set_progress", ex.handle, None, None)
    out = []
    for i, p in enumerate(arr):
        Rr, Tr = keep[i][4], keep[i][5]
        rho = sanitize(Hops({tuple(int(x) for x in Rr[:, r]): Tr[:, :, r] for r in range(Rr.shape[1])}))
        E, mu, eb, eH, eF, ek, res, nit = list(p.result)
        info = {"mu": mu, "epsband": eb, "epsH": eH, "epsF": eF, "epskin": ek, "iterations": int(nit)}
        out.append((rho, E, info, bool(p.converged), res))
    return out
