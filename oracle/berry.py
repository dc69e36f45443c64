"""FP64 CPU oracle of the band-topology functions -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` may import this module, as the checker; the product never calls it.
Restates Spectrum/berry.jl in numpy on eigenvectors from the C oracle (LAPACK zheevr):

* ``statesgrid``  -- eigenvectors on range(0, 1, NX) x range(0, 1, NY) shifted by 1.34e-8
  (berry.jl:53-84), array [NX, NY, N, M];
* ``link``        -- L(S1, S2) = det(S1^H S2) / |det| (berry.jl:116-119);
* ``berry``       -- plaquette phases real(log(L(S00,S10) L(S10,S11) / L(S01,S11) / L(S00,S01))
  / 2 pi i) over plaquettes (i, j), i < NX - 1, j < NY - 1 (berry.jl:121-152);
* ``getcherns``   -- sum(berry(states[..., [n]])) per band (berry.jl:425-429);
* ``wilson1d``    -- product of the polar (SVD) factors of the consecutive overlaps around
  a periodic 1D grid, its eigenphases / 2 pi ascending (berry.jl:362-383).
"""
from __future__ import annotations

import numpy as np

from . import oracle as _o


def statesgrid(R, T, NX, NY=0, bandindices=None):
    NY = NX if NY < 1 else NY
    xs = np.linspace(0.0, 1.0, NX)
    ys = np.linspace(0.0, 1.0, NY)
    ks = np.array([[x, y] for y in ys for x in xs]).T + 1.34e-8  # first index fastest
    _, V = _o.bands(R, T, np.asfortranarray(ks), vectors=True)
    N = V.shape[0]
    bi = list(range(N)) if not bandindices else list(bandindices)
    out = np.zeros((NX, NY, N, len(bi)), dtype=np.complex128)
    for j in range(NY):
        for i in range(NX):
            out[i, j] = V[:, bi, i + NX * j]
    return out


def link(S1, S2):
    d = np.linalg.det(S1.conj().T @ S2)
    return d / abs(d)


def berry(states):
    n, m = states.shape[:2]
    F = np.zeros((n - 1, m - 1))
    for i in range(n - 1):
        for j in range(m - 1):
            S00, S10, S01, S11 = states[i, j], states[i + 1, j], states[i, j + 1], states[i + 1, j + 1]
            z = link(S00, S10) * link(S10, S11) / link(S01, S11) / link(S00, S01)
            F[i, j] = np.angle(z) / (2 * np.pi)
    return F


def getcherns(R, T, NX, NY=0, bands=None):
    g = statesgrid(R, T, NX, NY, bands)
    return [float(np.sum(berry(g[:, :, :, [i]]))) for i in range(g.shape[3])]


def wilson1d(states, j0=0):
    """states [NY, N, M] on a periodic 1D grid."""
    NY, _, M = states.shape
    F = np.eye(M, dtype=np.complex128)
    for j in range(1, NY + 1):
        a = states[(j0 + j - 1) % NY]
        b = states[(j0 + j) % NY]
        u, _, vt = np.linalg.svd(a.conj().T @ b)
        F = F @ (u @ vt)
    ev = np.linalg.eigvals(F)
    return np.sort(np.angle(ev) / (2 * np.pi))
