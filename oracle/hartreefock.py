"""FP64 CPU oracle of the Hartree-Fock self-consistency -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` and the ``cpu_baseline`` leg of ``bench.py`` may import this module,
as the checker / the timed CPU baseline.  The product library never calls it.

Restates, in numpy on dicts of N x N blocks (key -> matrix):

* ``meanfield``  -- HartreeFock update and energies (Meanfield/hartreefock.jl:184-289,
  :392-410): H_MF = 0 + h, then H_MF[L] -= v_L .* rho_L (Fock, no conjugate, keys of v
  that rho has), then H_MF[0][i, i] += (V0 diag(rho[0]))[i] with V0 = sum_L v_L;
  eps_H = 1/2 n^T V0 n, eps_F = -1/2 sum_L sum rho_L conj(rho_L) v_L.
* ``scf``        -- solvehartreefock with linear mixing (Meanfield/selfconsistent.jl:256-336,
  :366-410; Meanfield/fixedpoint.jl:39-194): per iteration H_MF[rho0], mu from an
  eigenvalues pass (chemicalpotential), rho1 = getdensitymatrix(H_MF, mu), E = eps_band -
  (eps_H + eps_F), x1 <- beta x1 + (1 - beta) x0, residual ||x1 - x0|| / ||x1|| (2-norm
  of the per-block Frobenius norms), x0 <- x1, stop at residual < tol; then the energies
  are refreshed at the final rho and eps_kin = eps_band - 2 (eps_H + eps_F).

The k passes use the C oracle (LAPACK zheevr per k) through ``oracle.oracle``.
"""
from __future__ import annotations

import numpy as np

from . import oracle as _o


def _zero(d):
    return (0,) * d


def meanfield(h: dict, v: dict, rho: dict, hartree: bool = True, fock: bool = True):
    """(H_MF dict, eps_H, eps_F) with H_MF keyed by keys(h), then keys(v) absent from h,
    then the zero key."""
    d = len(next(iter(h)))
    N = next(iter(h.values())).shape[0]
    keys = list(h) + [L for L in v if L not in h]
    zk = _zero(d)
    if zk not in keys:
        keys.append(zk)
    hmf = {L: np.zeros((N, N), dtype=np.complex128) for L in keys}
    for L, t in h.items():
        hmf[L] = hmf[L] + t
    if fock:
        for L, vL in v.items():
            if L in rho:
                hmf[L] = hmf[L] - vL * rho[L]
    V0 = sum(v.values()) if v else np.zeros((N, N), dtype=np.complex128)
    eH = eF = 0.0
    if hartree:
        n = np.diag(rho[zk]).copy()
        dvec = V0 @ n
        for i in range(N):
            hmf[zk][i, i] += dvec[i]
        eH = float(np.real(0.5 * (n @ V0 @ n)))
    if fock:
        eF = float(np.real(-0.5 * sum(np.sum(rho[L] * np.conj(rho[L]) * vL) for L, vL in v.items()
                                      if L in rho)))
    return hmf, eH, eF


def _pack(blocks: dict):
    keys = list(blocks)
    d = len(keys[0])
    N = next(iter(blocks.values())).shape[0]
    R = np.asfortranarray(np.array(keys, dtype=np.int32).T.reshape(d, len(keys)))
    T = np.empty((N, N, len(keys)), dtype=np.complex128, order="F")
    for r, k in enumerate(keys):
        T[:, :, r] = blocks[k]
    return R, T


def scf(h: dict, v: dict, rho0: dict, filling: float, ks, T: float = 0.0, beta: float = 1.0,
        tol: float = 1e-7, iterations: int = 500, hartree: bool = True, fock: bool = True,
        kweights=None, nthreads=None):
    """Returns dict(rho, E, mu, eband, eH, eF, ekin, residual, iterations, converged,
    history (E, residual, mu) per iteration, hmf)."""
    x0 = {L: np.array(b, dtype=np.complex128) for L, b in rho0.items()}
    keys = list(x0)
    Rr = np.asfortranarray(np.array(keys, dtype=np.int32).T.reshape(len(keys[0]), len(keys)))
    history = []
    residual, eband, mu = np.inf, 0.0, 0.0
    converged = False
    it = 0
    while it < iterations:
        it += 1
        hmf, eH, eF = meanfield(h, v, x0, hartree, fock)
        R, Tm = _pack(hmf)
        b = _o.bands(R, Tm, ks, nthreads=nthreads)
        mu = _o.chemicalpotential(b, filling, T, kweights)
        rho, eband = _o.densitymatrix(R, Tm, ks, Rr, mu=mu, Temp=T, kweights=kweights,
                                      nthreads=nthreads)
        E = eband - (eH + eF)
        x1 = {L: beta * rho[:, :, r] + (1 - beta) * x0[L] for r, L in enumerate(keys)}
        dn = np.linalg.norm([np.linalg.norm(x1[L] - x0[L]) for L in keys])
        xn = np.linalg.norm([np.linalg.norm(x1[L]) for L in keys])
        residual = dn / xn if xn > np.finfo(float).eps else dn
        x0 = x1
        history.append((E, residual, mu))
        if residual < tol:
            converged = True
            break
    hmf, eH, eF = meanfield(h, v, x0, hartree, fock)
    dc = eH + eF
    return dict(rho=x0, E=eband - dc, mu=mu, eband=eband, eH=eH, eF=eF, ekin=eband - 2 * dc,
                residual=residual, iterations=it, converged=converged, history=np.array(history),
                hmf=hmf)
