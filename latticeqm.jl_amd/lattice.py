"""Lattice geometry and real-space hopping generators (host set-up, run once per model).

Restates the parts of LatticeQM's ``Structure`` and ``Operators`` modules that build
the exact ``Hops`` the band path is fed:

* ``Lattice``: Bravais basis A (3 x d, spatial rows padded to 3) and cartesian orbital
  positions (3 x n) plus extra labels such as ``sublattice`` (Structure/Lattices/type.jl);
* ``honeycomb`` (Structure/Geometries/2d.jl:50-52, :88-94);
* ``getneighborcells`` (Structure/Lattices/neighbors.jl:56-150): cells up to the k-th
  shell found in a Lagrange-reduced basis, half-space representatives;
* ``gethops`` (Operators/gethops.jl:33-140): block (i, j) of key L is
  t(r_i + A L, r_j), entries below sqrt(eps) dropped, the -L block is the adjoint,
  all-zero blocks removed (trim!, TightBinding/types.jl:278-294);
* ``twist`` / ``honeycomb_twisted`` (Structure/Lattices/twist.jl:45-173,
  Structure/Geometries/2d.jl:180-227): commensurate twisted bilayer on the moire
  supercell S = [n, -n-m; n+m, 2n+m], static layer at +z/2, the other layer parity
  flipped to -z/2, tiled 3 x 3, rotated by the twist angle and cropped to the cell
  (supercell.jl:32-61, :171-268);
* ``nearestneighbor`` (Operators/nearestneighbor.jl:22-62), ``t_graphene`` /
  ``graphene`` (Operators/graphene.jl:17-28, :228-273), ``gethubbard``
  (Operators/interactions.jl:8-28).
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np

from .hops import Hops
from .models import addspin

DEFAULT_PRECISION = float(np.sqrt(np.finfo(float).eps))

_A_TRI = np.array([[np.cos(np.pi / 6), np.cos(np.pi / 6)],
                   [-np.sin(np.pi / 6), np.sin(np.pi / 6)],
                   [0.0, 0.0]])


@dataclass
class Lattice:
    """A: 3 x d lattice vectors (columns); positions: 3 x n cartesian orbital positions."""
    A: np.ndarray
    positions: np.ndarray
    extra: dict = field(default_factory=dict)

    @property
    def d(self) -> int:
        return self.A.shape[1]

    @property
    def n(self) -> int:
        return self.positions.shape[1]

    def label(self, name: str) -> np.ndarray:
        return np.asarray(self.extra[name])


def honeycomb(a: float = 1.0) -> Lattice:
    """Two-site honeycomb, nearest-neighbour distance a: A = a sqrt(3) A_tri, sites at
    0 and (a1 + a2)/3, sublattice labels 0 (A) and 1 (B)."""
    A = a * np.sqrt(3.0) * _A_TRI
    frac = np.array([[0.0, 1.0 / 3.0], [0.0, 1.0 / 3.0]])
    return Lattice(A, A @ frac, {"sublattice": np.array([0, 1])})


def _lattice_reduce(A: np.ndarray):
    """Lagrange reduction of a 2D basis; returns (reduced basis, integer transform U)."""
    d = A.shape[1]
    U = np.eye(d, dtype=np.int64)
    if d != 2:
        return A.astype(float), U
    b1, b2 = A[:, 0].astype(float), A[:, 1].astype(float)
    for _ in range(1000):
        if np.linalg.norm(b1) > np.linalg.norm(b2):
            b1, b2 = b2, b1
            U = U[:, ::-1].copy()
        m = int(np.round(np.dot(b1, b2) / np.dot(b1, b1)))
        if m == 0:
            break
        b2 = b2 - m * b1
        U[:, 1] -= m * U[:, 0]
    return np.stack([b1, b2], axis=1), U


def getneighborcells(A: np.ndarray, k: int = 1, halfspace: bool = True, innerpoints: bool = False,
                     excludeorigin: bool = True):
    """Integer cell offsets of the k-th neighbour shell (all shells up to k with
    innerpoints), one of each +-pair when halfspace."""
    d = A.shape[1]
    Ared, U = _lattice_reduce(A)
    n = int(np.ceil(np.sqrt(3.0) * (k + 1)))
    IJ = np.array(list(itertools.product(range(-n, n + 1), repeat=d))).T  # first index fastest?
    # Iterators.product varies the first coordinate fastest
    IJ = IJ[::-1, :] if d > 1 else IJ
    D = np.round(np.linalg.norm(Ared @ IJ, axis=0), 9)
    dist = np.unique(D)
    sel = D <= dist[k] if innerpoints else D == dist[k]
    cells = [tuple(int(x) for x in U @ IJ[:, c]) for c in np.nonzero(sel)[0]]
    if halfspace:
        seen, keep = set(), []
        for v in cells:
            if all(x == 0 for x in v):
                if not excludeorigin:
                    keep.append(v)
            elif v not in seen and tuple(-x for x in v) not in seen:
                keep.append(v)
                seen.add(v)
        return keep
    if excludeorigin:
        cells = [v for v in cells if any(cells_x != 0 for cells_x in v)]
    return cells


def _hoppingblock(lat: Lattice, L, t, precision: float) -> np.ndarray:
    shift = lat.A @ np.asarray(L, dtype=float)
    R = lat.positions
    if getattr(t, "vectorized", False):
        # t(dr) on the whole (3, n, n) array of r_i + A L - r_j
        dr = (R + shift[:, None])[:, :, None] - R[:, None, :]
        M = np.asarray(t(dr), dtype=np.complex128)
        M[np.abs(M) < precision] = 0.0
        return M
    M = np.zeros((lat.n, lat.n), dtype=np.complex128)
    for i in range(lat.n):
        ri = R[:, i] + shift
        for j in range(lat.n):
            val = t(ri, R[:, j])
            if abs(val) >= precision:
                M[i, j] = val
    return M


def gethops(lat: Lattice, t, neighbors=None, cellrange: int = 2,
            precision: float = DEFAULT_PRECISION) -> Hops:
    """Hops of the pair function t over the neighbour cells (all shells up to ``cellrange``,
    half space + origin) or over an explicit list ``neighbors``."""
    if neighbors is None:
        neighbors = getneighborcells(lat.A, cellrange, halfspace=True, innerpoints=True,
                                     excludeorigin=False)
    h = Hops()
    for L in neighbors:
        L = tuple(int(x) for x in L)
        M = _hoppingblock(lat, L, t, precision)
        h[L] = M
        if any(x != 0 for x in L):
            h[tuple(-x for x in L)] = M.conj().T
    for L in list(h.keys()):  # trim!
        if not np.any(h[L] != 0):
            del h.data[L]
    return h


def nearestneighbor(lat: Lattice, t0: float = -1.0, a: float = 1.0, halfwidth: float = 0.01,
                    cellrange: int = 2) -> Hops:
    """t0 between orbitals at distance within (a - halfwidth, a + halfwidth)."""
    def t(r1, r2):
        r = np.linalg.norm(r1[:3] - r2[:3])
        return t0 if (a + halfwidth > r > a - halfwidth) else 0.0
    return gethops(lat, t, cellrange=cellrange)


def t_graphene(r1, r2, tmin: float = 1e-5, tz: float = 0.46, t0: float = 1.0, l_inter: float = 0.125,
               l_intra: float = 0.08, z: float = 3.0, a: float = 1.0, dmin: float = 0.1,
               dmax: float = 5.0) -> float:
    """-t0 (1 - chi) e^{-|D - a|/l_intra} - tz chi e^{-|D - z|/l_inter}, chi = dz^2/D^2,
    zero outside [dmin, dmax] or below tmin in magnitude."""
    dr = np.asarray(r1[:3]) - np.asarray(r2[:3])
    D = float(np.sqrt(np.sum(dr * dr)))
    if dmax < D or D < dmin:
        return 0.0
    chi = dr[2] ** 2 / D ** 2
    v = -t0 * (1 - chi) * np.exp(-abs(D - a) / l_intra) - tz * chi * np.exp(-abs(D - z) / l_inter)
    return 0.0 if abs(v) < tmin else float(v)


def _t_graphene_vec(tmin: float = 1e-5, tz: float = 0.46, t0: float = 1.0, l_inter: float = 0.125,
                    l_intra: float = 0.08, z: float = 3.0, a: float = 1.0, dmin: float = 0.1,
                    dmax: float = 5.0):
    """t_graphene on an array of separation vectors dr (3, ...)."""
    def t(dr):
        D = np.sqrt(np.sum(dr * dr, axis=0))
        ok = (D >= dmin) & (D <= dmax)
        Ds = np.where(ok, D, 1.0)
        chi = dr[2] ** 2 / Ds ** 2
        v = -t0 * (1 - chi) * np.exp(-np.abs(Ds - a) / l_intra) - tz * chi * np.exp(-np.abs(Ds - z) / l_inter)
        return np.where(ok & (np.abs(v) >= tmin), v, 0.0)
    t.vectorized = True
    return t


def graphene(lat: Lattice, mode: str = "nospin", cellrange: int = 2, **kw) -> Hops:
    """Graphene Hamiltonian with the t_graphene hopping (keyword overrides in kw)."""
    h = gethops(lat, _t_graphene_vec(**kw), cellrange=cellrange, precision=0.0)
    return addspin(h, "spinhalf") if mode == "spinhalf" else h


# ---- commensurate twisted bilayers ------------------------------------------------
def twistangle(n: int, m: int = 1, degrees: bool = False) -> float:
    """cos a = (3n^2 + 3nm + m^2/2) / (3n^2 + 3nm + m^2)  (twist.jl:44-47)."""
    a = float(np.arccos((3.0 * n ** 2 + 3 * n * m + m ** 2 / 2.0) / (3.0 * n ** 2 + 3 * n * m + m ** 2)))
    return a * 180 / np.pi if degrees else a


def moire_supercell(n: int, m: int = 1) -> np.ndarray:
    """Columns: the moire vectors in the primitive basis, [n, -n-m; n+m, 2n+m]."""
    if np.gcd(n, m) != 1:
        raise ValueError("twist indices must be coprime")
    return np.array([[n, -n - m], [n + m, 2 * n + m]], dtype=np.int64)


def supercellpoints(S: np.ndarray, offset: float = float(np.sqrt(np.finfo(float).eps))) -> np.ndarray:
    """Integer points x with inv(S) x in [0, 1)^d (up to offset), (d, |det S|)."""
    d = S.shape[0]
    corners = S @ np.array(list(itertools.product([0, 1], repeat=d))).T
    lo, hi = corners.min(axis=1), corners.max(axis=1)
    Phi = np.linalg.inv(S.astype(float))
    pts = []
    for x in itertools.product(*[range(int(a), int(b) + 1) for a, b in zip(lo, hi)]):
        f = Phi @ np.array(x, dtype=float)
        if np.all((f < 1 - offset) & (f >= -offset)):
            pts.append(x)
    n = int(round(abs(np.linalg.det(S))))
    if len(pts) != n:
        raise RuntimeError("supercell enumeration failed")
    return np.array(pts, dtype=np.int64).T


def twist(lat: Lattice, n: int, z: float = 3.0, m: int = 1) -> Lattice:
    """Twisted bilayer of a 2D triangular-Bravais lattice (positions in the returned
    lattice are cartesian, folded into the moire cell; extra labels: layer, sublattice)."""
    A2 = lat.A[:2, :]
    alpha = twistangle(n, m)
    S = moire_supercell(n, m)
    Am = A2 @ S
    frac = np.linalg.solve(A2, lat.positions[:2])  # primitive fractional coordinates
    zs = lat.positions[2]
    sub = lat.extra.get("sublattice", np.zeros(lat.n))
    P = supercellpoints(S)
    off = float(np.sqrt(np.finfo(float).eps))
    xs, zz, layer, sl = [], [], [], []
    # static layer: translated to +z/2
    for p in P.T:
        for s in range(lat.n):
            xs.append(A2 @ (frac[:, s] + p))
            zz.append(zs[s] + z / 2)
            layer.append(0.0)
            sl.append(sub[s])
    nstatic = len(xs)
    # rotated layer: +z/2, full parity flip, fold, supercell, 3x3 tile, rotate, crop
    fflip = np.mod(-frac, 1.0)
    c, sn = np.cos(alpha), np.sin(alpha)
    Rot = np.array([[c, -sn], [sn, c]])
    for i, j in itertools.product((-1, 0, 1), repeat=2):
        tile = Am @ np.array([i, j], dtype=float)
        for p in P.T:
            for s in range(lat.n):
                r = Rot @ (A2 @ (fflip[:, s] + p) + tile)
                f = np.linalg.solve(Am, r)
                if np.all((f < 1 - off) & (f >= -off)):
                    xs.append(r)
                    zz.append(-(zs[s] + z / 2))
                    layer.append(1.0)
                    sl.append(sub[s])
    if len(xs) - nstatic != nstatic:
        raise RuntimeError("twisted layer lost or duplicated atoms at the cell boundary")
    pos = np.zeros((3, len(xs)))
    pos[:2] = np.array(xs).T
    pos[2] = zz
    A = np.zeros((3, 2))
    A[:2] = Am
    return Lattice(A, pos, {"layer": np.array(layer), "sublattice": np.array(sl)})


def honeycomb_twisted(N: int, a: float = 1.0, z: float = 3.0) -> Lattice:
    """Twisted bilayer graphene with twist index N: 4(3N^2 + 3N + 1) orbitals."""
    return twist(honeycomb(a), N, z=z, m=1)


def gethubbard(lat: Lattice, neighbors=((0, 0),), mode: str = "nospin", a: float = 0.5,
               U: float = 1.0) -> Hops:
    """U heaviside(a - |r_i + A L - r_j|) on the listed cells, then addspin(mode)."""
    def t(r1, r2):
        x = a - np.linalg.norm(r1[:3] - r2[:3])
        return U * (0.0 if x < 0 else (1.0 if x > 0 else 0.5))
    v = gethops(lat, t, neighbors=[tuple(L) for L in neighbors])
    spin = {"σx": "sigmax", "σ0": "spinhalf"}.get(mode, mode)
    return addspin(v, spin)


def _nnbondlength(lat: Lattice) -> float:
    """Smallest nonzero interatomic distance over the origin and first-shell cells."""
    cells = getneighborcells(lat.A, 1, halfspace=False, innerpoints=True, excludeorigin=False)
    R = lat.positions
    dmin = np.inf
    for c in cells:
        sh = lat.A @ np.asarray(c, dtype=float)
        d = np.linalg.norm((R + sh[:, None])[:, :, None] - R[:, None, :], axis=0)
        d = d[d > DEFAULT_PRECISION]
        if d.size:
            dmin = min(dmin, float(d.min()))
    return dmin


def haldane(lat: Lattice, t2: float = 0.1, phi: float = np.pi / 2, a0: float | None = None,
            cellrange: int = 1) -> Hops:
    """The Haldane next-nearest-neighbour term (Operators/haldane.jl:47-86): for every pair of
    orbitals at distance sqrt(3) a0 (a0 the nearest-neighbour bond), through their common
    nearest neighbour R0, block (i, j) of key L gains t2 exp(i phi sign(cross(-y, x))) with
    x = r_i + A L - R0, y = r_j - R0 (cross(a, b) = a_1 b_2 - a_2 b_1: the chirality)."""
    a0 = _nnbondlength(lat) if a0 is None else a0
    cells = getneighborcells(lat.A, cellrange, halfspace=False, innerpoints=True, excludeorigin=False)
    near = getneighborcells(lat.A, 1, halfspace=False, innerpoints=True, excludeorigin=False)
    R = lat.positions
    R0s = np.concatenate([R + (lat.A @ np.asarray(c, dtype=float))[:, None] for c in near], axis=1)
    h = Hops()
    d3 = np.sqrt(3.0) * a0
    for L in cells:
        sh = lat.A @ np.asarray(L, dtype=float)
        M = np.zeros((lat.n, lat.n), dtype=np.complex128)
        for i in range(lat.n):
            Ri = R[:, i] + sh
            for j in range(lat.n):
                Rj = R[:, j]
                if not np.isclose(np.linalg.norm(Ri - Rj), d3, rtol=1e-6):
                    continue
                for q in range(R0s.shape[1]):
                    x = Ri[:2] - R0s[:2, q]
                    y = Rj[:2] - R0s[:2, q]
                    if np.isclose(np.linalg.norm(x), a0, rtol=1e-6) and np.isclose(np.linalg.norm(y), a0, rtol=1e-6):
                        cr = (-y[0]) * x[1] - (-y[1]) * x[0]
                        M[i, j] += t2 * np.exp(1j * phi * np.sign(cr))
                        break
        if np.any(M != 0):
            h[tuple(int(v) for v in L)] = M
    return h


def getcurrentoperators(lat: Lattice, hops, **kw) -> list:
    """Current operators J_a = dH(k)/dk_a, one Hops per lattice direction a = 0 .. d-1
    (Operators/current.jl:20-50).

    From a Hops: block (i, j) of key L times -i (r_i + A L - r_j)_a, the factor applied to
    n x n blocks per orbital pair when the Hops carries an internal multiplicity
    n = D / (number of orbitals) (spin: kron(t, sigma) blocks).  From a pair function
    t(r1, r2): the Hops of -i (r1 - r2)_a t(r1, r2) (keywords go to gethops)."""
    if callable(hops) and not isinstance(hops, Hops):
        t = hops

        def ta(a):
            return lambda r1, r2: -1j * (r1 - r2)[a] * t(r1, r2)

        return [gethops(lat, ta(a), **kw) for a in range(lat.d)]
    N, D = lat.n, hops.N
    if not (D >= N and D % N == 0):
        raise AssertionError("Hops dimension must be a multiple of the number of orbitals")
    n = D // N
    r = lat.positions
    ones = np.ones((n, n))
    Js = []
    for a in range(lat.d):
        J = Hops()
        for L, t in hops.items():
            r0 = r[a] + (lat.A @ np.asarray(L, dtype=float))[a]
            J[L] = t * np.kron(-1j * (r0[:, None] - r[a][None, :]), ones)
        Js.append(J)
    return Js


def addhops(h: Hops, extra: Hops) -> Hops:
    """h + extra key by key (addhops!, TightBinding/types.jl)."""
    out = h.copy()
    for R, t in extra.items():
        out[R] = out[R] + t if R in out else t
    return out
