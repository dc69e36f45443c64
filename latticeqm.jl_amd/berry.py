"""Band topology on GPU eigenvector grids, mirroring LatticeQM's Spectrum/berry.jl.

=====================  ======================================================
this module            reference (Spectrum/berry.jl)
=====================  ======================================================
``statesgrid``         statesgrid (:53-84)
``statesgrid1D``       statesgrid1D (:86-106)
``berry``              berry(statesgrid) / berry(H, NX, NY, bandindices) (:121-160)
``getcherns``          getcherns (:425-429)
``getwindnum``         getwindnum (:441-443)
``spectralgaps``       spectralgaps (:473-493)
``gapcherns``          gapcherns (:522-540)
``wilson1d``           Wilson1D (:343-383)
``wilsonslice1d``      WilsonSlice1D (:395-423)
``wilson2d``           Wilson2D (:286-297)
``berryalongpath``     berryalongpath (:168-231)
=====================  ======================================================

Band indices are 0-based here (the reference's band n is n - 1).  The eigenvectors come from
the GPU eigensolver (``api.eigen``); the link variables L(S1, S2) = det(S1^H S2) / |det| of
every plaquette edge are formed on the device in one call (``lqm_link_phases``: the overlaps
on the matrix cores, the determinant phases by LU), and the Wilson-loop overlaps likewise
(``lqm_link_overlaps``); the host only multiplies the unit phases around each plaquette and,
for Wilson loops, the small M x M polar factors.
"""
from __future__ import annotations

import ctypes as _c

import numpy as np

from . import _lib
from .hops import Hops


def _exec(executor):
    from .api import to_executor

    return to_executor(executor or "gpu")


def _eigvecs(H, ks, executor):
    from .api import eigen

    return eigen(H, np.asfortranarray(ks), executor=executor)


def _pack_states(blocks):
    """[npts, N, M] complex -> Fortran N x M x npts buffer of the C ABI."""
    return np.asfortranarray(np.transpose(np.asarray(blocks, dtype=np.complex128), (1, 2, 0)))


def link_phases(blocks, pairs, executor=None) -> np.ndarray:
    """Unit phases det(S_a^H S_b) / |det| for each (a, b) in pairs; blocks [npts, N, M]."""
    U = _pack_states(blocks)
    N, M, npts = U.shape
    P = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
    out = np.zeros(2 * len(P))
    if len(P):
        _lib.call("lqm_link_phases", _exec(executor).handle, N, M, npts, _c.c_void_p(U.ctypes.data), len(P),
                  _c.c_void_p(P.ctypes.data), _c.c_void_p(out.ctypes.data))
    return out[0::2] + 1j * out[1::2]


def link_overlaps(blocks, pairs, executor=None) -> np.ndarray:
    """S_a^H S_b [nlink, M, M] for each (a, b) in pairs; blocks [npts, N, M]."""
    U = _pack_states(blocks)
    N, M, npts = U.shape
    P = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
    out = np.zeros((M, M, max(len(P), 1)), dtype=np.complex128, order="F")
    if len(P):
        _lib.call("lqm_link_overlaps", _exec(executor).handle, N, M, npts, _c.c_void_p(U.ctypes.data), len(P),
                  _c.c_void_p(P.ctypes.data), _c.c_void_p(out.ctypes.data))
    return np.transpose(out[:, :, :len(P)], (2, 0, 1))


def statesgrid(H, NX: int, NY: int = 0, bandindices=(), executor=None):
    """Eigenvectors on range(0, 1, NX) x range(0, 1, NY) (shifted by 1.34e-8): returns
    (kgrid [NX, NY, 2], midkgrid [NX-1, NY-1, 2], states [NX, NY, N, M])."""
    NY = NX if NY < 1 else NY
    xs, ys = np.linspace(0.0, 1.0, NX), np.linspace(0.0, 1.0, NY)
    kgrid = np.stack(np.meshgrid(xs, ys, indexing="ij"), axis=-1)
    mx = np.linspace(0.5 / (NX - 1), 1 - 0.5 / (NX - 1), NX - 1) if NX > 1 else np.zeros(0)
    my = np.linspace(0.5 / (NY - 1), 1 - 0.5 / (NY - 1), NY - 1) if NY > 1 else np.zeros(0)
    midkgrid = np.stack(np.meshgrid(mx, my, indexing="ij"), axis=-1)
    ks = kgrid.reshape(-1, 2, order="F").T + 1.34e-8  # first index fastest
    _, V = _eigvecs(H, ks, executor)
    N = V.shape[0]
    bi = list(range(N)) if len(bandindices) == 0 else [int(b) for b in bandindices]
    states = np.transpose(V[:, bi, :].reshape(N, len(bi), NX, NY, order="F"), (2, 3, 0, 1))
    return kgrid, midkgrid, np.ascontiguousarray(states)


def statesgrid1D(H, NX: int, bandindices=(), executor=None):
    """Eigenvectors on range(0, 1, NX) of a one-dimensional model: (kgrid, states [NX, N, M])."""
    kgrid = np.linspace(0.0, 1.0, NX)
    _, V = _eigvecs(H, kgrid.reshape(1, -1), executor)
    N = V.shape[0]
    bi = list(range(N)) if len(bandindices) == 0 else [int(b) for b in bandindices]
    return kgrid, np.ascontiguousarray(np.transpose(V[:, bi, :], (2, 0, 1)))


def _plaquettes(states, executor):
    n, m, N, M = states.shape
    blocks = states.reshape(n * m, N, M, order="F")  # point p = i + n j
    p = lambda i, j: i + n * j
    xl = [(p(i, j), p(i + 1, j)) for j in range(m) for i in range(n - 1)]
    yl = [(p(i, j), p(i, j + 1)) for j in range(m - 1) for i in range(n)]
    ph = link_phases(blocks, xl + yl, executor)
    LX = ph[:len(xl)].reshape(n - 1, m, order="F")
    LY = ph[len(xl):].reshape(n, m - 1, order="F")
    return LX, LY


def berry(states_or_H, NX: int | None = None, NY: int = 0, bandindices=(0,), executor=None):
    """Plaquette phases F[i, j] of a states grid [NX, NY, N, M] (the non-abelian flux of the M
    columns through plaquette (i, j)-(i+1, j+1)); or, for a Hamiltonian, (midkgrid, F) on the
    NX x NY grid of ``bandindices``."""
    if NX is not None:
        _, mid, st = statesgrid(states_or_H, NX, NY, bandindices, executor)
        return mid, berry(st, executor=executor)
    states = np.asarray(states_or_H)
    LX, LY = _plaquettes(states, executor)
    # S00 -> S10 -> S11 <- S01 <- S00: L(S00,S10) L(S10,S11) L(S01,S11)^-1 L(S00,S01)^-1
    z = LX[:, :-1] * LY[1:, :] * np.conj(LX[:, 1:]) * np.conj(LY[:-1, :])
    return np.angle(z) / (2 * np.pi)


def getcherns(H, NX: int, NY: int = 0, bands=(), executor=None):
    """Chern number of each band in ``bands`` (all when empty): the summed abelian plaquette
    fluxes on the NX x NY grid.  All bands' links in one device call."""
    _, _, g = statesgrid(H, NX, NY, bands, executor)
    n, m, N, M = g.shape
    # every band as its own one-column state set: point p of band b is block b n m + p
    blocks = np.concatenate([g[:, :, :, [b]].reshape(n * m, N, 1, order="F") for b in range(M)], axis=0)
    p = lambda i, j: i + n * j
    xl = [(p(i, j), p(i + 1, j)) for j in range(m) for i in range(n - 1)]
    yl = [(p(i, j), p(i, j + 1)) for j in range(m - 1) for i in range(n)]
    base = np.array(xl + yl, dtype=np.int64)
    pairs = np.concatenate([base + b * n * m for b in range(M)], axis=0)
    ph = link_phases(blocks, pairs, executor).reshape(M, -1)
    out = []
    for b in range(M):
        LX = ph[b, :len(xl)].reshape(n - 1, m, order="F")
        LY = ph[b, len(xl):].reshape(n, m - 1, order="F")
        z = LX[:, :-1] * LY[1:, :] * np.conj(LX[:, 1:]) * np.conj(LY[:-1, :])
        out.append(float(np.sum(np.angle(z) / (2 * np.pi))))
    return out


def getwindnum(H, NX: int, NY: int = 0, bandnr: int = 0, executor=None) -> float:
    """Sum of the Chern numbers of bands 0 .. bandnr (the winding number of band bandnr)."""
    return float(sum(getcherns(H, NX, NY, list(range(bandnr + 1)), executor)))


def spectralgaps(H, NX: int, NY: int | None = None, gaptol: float = 1e-6, executor=None):
    """Global gaps from eigenvalues on the statesgrid k-grid: [(n, elo, ehi)] with n bands below
    a gap [elo, ehi] wider than gaptol."""
    from .api import bandmatrix

    NY = NX if NY is None else NY
    xs, ys = np.linspace(0.0, 1.0, NX), np.linspace(0.0, 1.0, NY)
    ks = np.stack(np.meshgrid(xs, ys, indexing="ij"), axis=-1).reshape(-1, 2, order="F").T + 1.34e-8
    E, _ = bandmatrix(H, np.asfortranarray(ks), executor=executor)
    emin, emax = E.min(axis=1), E.max(axis=1)
    return [(n + 1, float(emax[n]), float(emin[n + 1])) for n in range(E.shape[0] - 1)
            if emin[n + 1] - emax[n] > gaptol]


def gapcherns(H, NX: int, NY: int | None = None, gaptol: float = 1e-6, executor=None):
    """Chern number of each spectral gap: the non-abelian flux of the manifold of the n bands
    below it, rounded: [(n, elo, ehi, chern)]."""
    NY = NX if NY is None else NY
    gaps = spectralgaps(H, NX, NY, gaptol, executor)
    if not gaps:
        return []
    _, _, st = statesgrid(H, NX, NY, (), executor)
    return [(n, lo, hi, int(round(float(np.sum(berry(st[:, :, :, :n], executor=executor))))))
            for (n, lo, hi) in gaps]


def _polar(O):
    u, _, vt = np.linalg.svd(O)
    return u @ vt


def wilson1d(states_or_H, arg=None, bandindices=(), executor=None):
    """Wilson loop of a periodic 1D states grid [NY, N, M] (start j0 = ``arg``, default 0):
    (eigenphases / 2 pi ascending, eigenvectors); the consecutive overlaps come from the
    device in one call, their polar factors are multiplied on the host.  For a Hamiltonian
    (1D), ``arg`` is NX: returns (phases, vectors, states)."""
    if isinstance(states_or_H, Hops) or hasattr(states_or_H, "pack"):
        _, st = statesgrid1D(states_or_H, int(arg), bandindices, executor)
        en, U = wilson1d(st, 0, executor=executor)
        return en, U, st
    st = np.asarray(states_or_H)
    j0 = 0 if arg is None else int(arg)
    NY = st.shape[0]
    pairs = [((j0 + j - 1) % NY, (j0 + j) % NY) for j in range(1, NY + 1)]
    O = link_overlaps(st, pairs, executor)
    F = np.eye(st.shape[2], dtype=np.complex128)
    for Ol in O:
        F = F @ _polar(Ol)
    ev, U = np.linalg.eig(F)
    en = np.angle(ev) / (2 * np.pi)
    perm = np.argsort(en, kind="stable")
    return en[perm], U[:, perm]


def wilsonslice1d(states, dim: int = 0, executor=None):
    """Wilson loops along axis 1 - dim of a [NX, NY, N, M] grid for every slice of axis dim."""
    st = np.asarray(states)
    if dim == 1:
        st = np.transpose(st, (1, 0, 2, 3))
    en, U = [], []
    for i in range(st.shape[0]):
        e, u = wilson1d(st[i], 0, executor=executor)
        en.append(e)
        U.append(u)
    return np.array(en), np.array(U)


def wilson2d(H, NX: int, NY: int = 0, bandindices=(0,), executor=None):
    """Wilson loops of the (NX-1) x (NY-1) periodic grid along both axes: (en1, U1, en2, U2)."""
    NY = NX if NY < 1 else NY
    _, _, st = statesgrid(H, NX, NY, bandindices, executor)
    g = st[: NX - 1, : NY - 1]
    en1, U1 = wilsonslice1d(g, 0, executor)
    en2, U2 = wilsonslice1d(g, 1, executor)
    return en1, U1, en2, U2


def _rot(theta):
    return np.array([[np.cos(theta), -np.sin(theta)], [np.sin(theta), np.cos(theta)]])


def berryalongpath(H, kpoints, executor=None) -> np.ndarray:
    """Abelian Berry curvature of every band along a path: per path point, the plaquette of the
    four points half-way to its neighbours (rotated by +-pi/2), [N, npath]."""
    kp = np.asarray(kpoints, dtype=np.float64)
    Np = kp.shape[1]
    k0 = kp[:, 0] - (kp[:, 1] - kp[:, 0])
    kN = kp[:, -1] + (kp[:, -1] - kp[:, -2])
    kp = np.column_stack([k0, kp, kN])
    kp = kp + 1.34e-6 * np.linalg.norm(kp[:, 1] - kp[:, 0]) * np.array([1.0, 0.7071067811865476])[:, None]
    corners = []
    for j in range(Np):
        kc = kp[:, j + 1]
        dR = (kp[:, j + 2] - kc) / 2
        dL = (kp[:, j] - kc) / 2
        dU = (_rot(np.pi / 2) @ dR + _rot(-np.pi / 2) @ dL) / 2
        dD = (_rot(-np.pi / 2) @ dR + _rot(np.pi / 2) @ dL) / 2
        corners += [kc + dR, kc + dU, kc + dD, kc + dL]
    _, V = _eigvecs(H, np.array(corners).T, executor)
    N = V.shape[0]
    blocks = np.transpose(V, (2, 0, 1))  # [4 Np, N, N]
    # per point: links (R, U), (U, L), (D, L), (R, D); band n's link is the diagonal entry n
    pairs = []
    for j in range(Np):
        r, u, d, l = 4 * j, 4 * j + 1, 4 * j + 2, 4 * j + 3
        pairs += [(r, u), (u, l), (d, l), (r, d)]
    O = link_overlaps(blocks, pairs, executor)
    dg = np.diagonal(O, axis1=1, axis2=2).reshape(Np, 4, N)
    ph = dg / np.abs(dg)
    z = ph[:, 0] * ph[:, 1] / ph[:, 2] / ph[:, 3]
    return (np.angle(z) / (2 * np.pi)).T
