"""FP64 CPU oracle of the eigenvector consumers -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` may import this module, as the checker; the product never calls it.
Restates, per k-point in numpy on eigenvectors from the C oracle (LAPACK zheevr):

* ``ldos``     -- Spectrum/dos.jl:57-90: n_i += Im(|psi_i|^2 / (w + iG - e)) per state and
  frequency, / Nw per k-point (the one-frequency result of the reference; for a window the
  frequency mean, see the C-ABI note), then * (-1 / (Nk pi));
* ``density``  -- dos.jl:20-52: n_i += |psi_i|^2 for e <= mu, / Nk;
* ``kubo``     -- LinearResponse/kubo.jl:11-31 inside opticalconductivity.jl:101-127:
  M = (U^H A U) .* transpose(U^H B U), A = J1(k), B = J2(k)^H, f = fermidirac(e; mu, T);
  out += (f_i - f_j) M_ij / ((e_i - e_j) - (w + iG)) over i != j; then * (-i / Nk).

k-weights, where given, replace the 1/Nk.
"""
from __future__ import annotations

import numpy as np

from . import oracle as _o


def _eig(R, T, ks):
    b, V = _o.bands(R, T, np.asfortranarray(ks), vectors=True)
    return b, V  # b [N, Nk], V [N, N, Nk] (column n = state n)


def ldos(R, T, ks, omegas, gamma, kweights=None):
    b, V = _eig(R, T, ks)
    om = np.atleast_1d(np.asarray(omegas, dtype=float))
    Nk = b.shape[1]
    n = np.zeros(b.shape[0])
    for k in range(Nk):
        g = np.zeros(b.shape[0])
        for w in om:
            g += np.imag(1.0 / (w + 1j * gamma - b[:, k]))
        g /= om.size
        wk = 1.0 if kweights is None else kweights[k]
        n += wk * (np.abs(V[:, :, k]) ** 2) @ g
    return n * (-1.0 / np.pi) / (Nk if kweights is None else 1.0)


def density(R, T, ks, mu, kweights=None):
    b, V = _eig(R, T, ks)
    Nk = b.shape[1]
    n = np.zeros(b.shape[0])
    for k in range(Nk):
        wk = 1.0 if kweights is None else kweights[k]
        n += wk * (np.abs(V[:, :, k]) ** 2) @ (b[:, k] <= mu).astype(float)
    return n / (Nk if kweights is None else 1.0)


def _bloch_all(R, T, ks):
    """J(k) = sum_R T_R e^{2 pi i k.R} for every k-point at once: [Nk, N, N]."""
    ph = np.exp(2j * np.pi * (np.asarray(ks, dtype=float).T @ np.asarray(R, dtype=float)))  # [Nk, nR]
    return np.einsum("kr,ijr->kij", ph, np.asarray(T))


def kubo(R, T, RJ1, TJ1, RJ2, TJ2, ks, omegas, mu=0.0, gamma=0.025, Temp=0.1, kweights=None, block=256):
    b, V = _eig(R, T, ks)
    om = np.atleast_1d(np.asarray(omegas, dtype=float))
    N, Nk = b.shape
    out = np.zeros(om.size, dtype=np.complex128)
    off = ~np.eye(N, dtype=bool)
    z = om[:, None] + 1j * gamma
    for k0 in range(0, Nk, block):
        sl = slice(k0, min(Nk, k0 + block))
        A = _bloch_all(RJ1, TJ1, ks[:, sl])
        B = np.conj(np.transpose(_bloch_all(RJ2, TJ2, ks[:, sl]), (0, 2, 1)))  # J2(k)^H
        U = np.transpose(V[:, :, sl], (2, 0, 1))
        Uh = np.conj(np.transpose(U, (0, 2, 1)))
        M = (Uh @ A @ U) * np.transpose(Uh @ B @ U, (0, 2, 1))
        for j, k in enumerate(range(sl.start, sl.stop)):
            fd = _o.fermidirac(b[:, k], mu, Temp)
            dF = (fd[:, None] - fd[None, :])[off]
            dE = (b[:, k][:, None] - b[:, k][None, :])[off]
            wk = 1.0 if kweights is None else kweights[k]
            out += wk * np.sum((dF * M[j][off])[None, :] / (dE[None, :] - z), axis=1)
    return out * (-1j) / (Nk if kweights is None else 1.0)
