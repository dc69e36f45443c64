"""Host-side containers of the band path: the Hops operator and k-grids.

``Hops`` mirrors ``TightBinding.Hops`` (TightBinding/types.jl:54-60): a mapping
from an integer lattice vector R to an N x N complex block t_R, with
H(k) = sum_R t_R exp(2 pi i k.R) (types.jl:8, :18-25).  ``pack`` produces the
flat column-major arrays the C ABI takes; keys keep insertion order, which is
the summation order on the device.

``regulargrid`` / ``Mesh`` mirror Structure/grid.jl:2-54.
"""
from __future__ import annotations

from typing import Dict, Iterable, Tuple

import numpy as np

Key = Tuple[int, ...]


def _key(R) -> Key:
    return tuple(int(x) for x in np.asarray(R).ravel())


class Hops:
    """R (tuple of ints) -> N x N complex128 block."""

    def __init__(self, data: Dict | Iterable | None = None):
        self.data: Dict[Key, np.ndarray] = {}
        if data is None:
            return
        items = data.items() if isinstance(data, dict) else data
        for R, t in items:
            self[R] = t

    # -- mapping interface -------------------------------------------------------
    def __getitem__(self, R) -> np.ndarray:
        return self.data[_key(R)]

    def __setitem__(self, R, t) -> None:
        t = np.array(t, dtype=np.complex128, copy=True)
        if t.ndim != 2 or t.shape[0] != t.shape[1]:
            raise ValueError("hopping blocks must be square matrices")
        self.data[_key(R)] = t

    def __contains__(self, R) -> bool:
        return _key(R) in self.data

    def __len__(self) -> int:
        return len(self.data)

    def __iter__(self):
        return iter(self.data.items())

    def keys(self):
        return self.data.keys()

    def values(self):
        return self.data.values()

    def items(self):
        return self.data.items()

    # -- shape ---------------------------------------------------------------------
    @property
    def N(self) -> int:
        """hopdim (types.jl:191-195)."""
        return next(iter(self.data.values())).shape[0]

    @property
    def d(self) -> int:
        return len(next(iter(self.data.keys())))

    def zerokey(self) -> Key:
        return (0,) * self.d

    # -- conversions ----------------------------------------------------------------
    def pack(self):
        """(R, T): R int32 (d, nR), T complex128 (N, N, nR), both Fortran order."""
        if not self.data:
            raise ValueError("empty Hops")
        keys = list(self.data.keys())
        N, d = self.N, self.d
        R = np.asfortranarray(np.array(keys, dtype=np.int32).T.reshape(d, len(keys)))
        T = np.empty((N, N, len(keys)), dtype=np.complex128, order="F")
        for r, k in enumerate(keys):
            blk = self.data[k]
            if blk.shape != (N, N):
                raise ValueError("all blocks must have the same size")
            T[:, :, r] = blk
        return R, T

    @classmethod
    def from_arrays(cls, R, T) -> "Hops":
        R = np.asarray(R)
        T = np.asarray(T)
        return cls({tuple(int(x) for x in R[:, r]): T[:, :, r] for r in range(R.shape[1])})

    def copy(self) -> "Hops":
        return Hops({k: v.copy() for k, v in self.data.items()})

    def zero(self) -> "Hops":
        """Same keys and shapes, zero blocks (types.jl:151-153)."""
        return Hops({k: np.zeros_like(v) for k, v in self.data.items()})

    def __call__(self, k) -> np.ndarray:
        """Dense H(k) on the host (fouriersum, types.jl:13-16)."""
        k = np.asarray(k, dtype=np.float64).ravel()
        out = np.zeros((self.N, self.N), dtype=np.complex128)
        for R, t in self.data.items():
            out += t * np.exp(1j * 2 * np.pi * float(np.dot(k, R)))
        return out

    def ishermitian(self, tol: float = float(np.sqrt(np.finfo(float).eps))) -> bool:
        """t_{-R} == t_R^H for every key (types.jl:165-173)."""
        for R, t in self.data.items():
            mR = tuple(-x for x in R)
            if mR not in self.data:
                return False
            if np.max(np.abs(self.data[mR] - t.conj().T), initial=0.0) > tol:
                return False
        return True

    def hermitianize(self) -> "Hops":
        """In place: t_R <- (t_R + t_{-R}^H)/2, t_{-R} <- t_R^H; a missing partner is
        added as the adjoint (types.jl:174-188)."""
        for R in list(self.data.keys()):
            mR = tuple(-x for x in R)
            if mR in self.data:
                self.data[R] = (self.data[R] + self.data[mR].conj().T) / 2
                self.data[mR] = self.data[R].conj().T.copy()
            else:
                self.data[mR] = self.data[R].conj().T.copy()
        return self


def regulargrid(nk: int = 100, dim: int = 2) -> np.ndarray:
    """Gamma-centred unshifted grid, first index fastest (Structure/grid.jl:42-54).

    Returns a (dim, n**dim) Fortran-ordered array with n = floor(nk**(1/dim) + 1e-9).
    """
    n = int(np.floor(nk ** (1.0 / dim) + 1e-9))
    it = np.arange(n, dtype=np.float64) / n
    mesh = np.meshgrid(*([it] * dim), indexing="ij")
    # Iterators.product varies the first coordinate fastest
    pts = np.stack([m.ravel(order="F") for m in mesh], axis=0)
    return np.asfortranarray(pts)


def kpath(corners, num_points: int = 60, B=None) -> np.ndarray:
    """Piecewise-linear path through the fractional k-points ``corners`` (d x m), about
    ``num_points`` points in all, spread over the segments in proportion to their length
    measured with the reciprocal basis B (identity: fractional lengths), both ends
    included (the getpath sampling behind Structure.kpath, Structure/Paths/path.jl:22-28)."""
    c = np.asarray(corners, dtype=np.float64)
    Bm = np.eye(c.shape[0]) if B is None else np.asarray(B, dtype=np.float64)
    seg = [np.linalg.norm(Bm @ (c[:, i + 1] - c[:, i])) for i in range(c.shape[1] - 1)]
    total = sum(seg) or 1.0
    pts = []
    for i, L in enumerate(seg):
        n = max(1, int(round((num_points - 1) * L / total)))
        t = np.arange(n) / n
        pts.append(c[:, i:i + 1] + t[None, :] * (c[:, i + 1] - c[:, i])[:, None])
    pts.append(c[:, -1:])
    return np.asfortranarray(np.concatenate(pts, axis=1))


class Mesh:
    """k-points with weights (Structure/grid.jl:2-20)."""

    def __init__(self, points, measure=None, multiplicity=None):
        self.points = np.asfortranarray(np.asarray(points, dtype=np.float64))
        L = self.points.shape[1]
        self.measure = np.full(L, 1.0 / L) if measure is None else np.asarray(measure, dtype=np.float64)
        self.multiplicity = np.ones(L, dtype=np.int64) if multiplicity is None else np.asarray(multiplicity)

    def weights(self) -> np.ndarray:
        """meshweights (grid.jl:20)."""
        return self.measure * self.multiplicity

    def __len__(self) -> int:
        return self.points.shape[1]


def points(ks) -> np.ndarray:
    """Structure.points (grid.jl:34-37): k-points as a (d, Nk) array."""
    if isinstance(ks, Mesh):
        return ks.points
    if hasattr(ks, "points"):
        return np.asfortranarray(np.asarray(ks.points, dtype=np.float64))
    a = np.asarray(ks, dtype=np.float64)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    return np.asfortranarray(a)
