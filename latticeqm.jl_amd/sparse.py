"""Sparse Hops and the partial-spectrum path (format = :sparse, num_bands).

=================================  ==========================================================
this module                        reference
=================================  ==========================================================
``SparseHops``                     Hops with sparse blocks, format = :sparse
                                   (TightBinding/types.jl:102)
``eigen_sparse(H, ks, num_bands,   Eigen.eigen_sparse(H(k); num_bands, sigma, which, tol,
sigma, which)``                    maxiter) (Eigen/eigen_sparse_krylovkit.jl:44-88) at every
                                   k-point: which = "LM" the num_bands eigenpairs nearest
                                   sigma, "SR" the smallest, "LR" the largest; ascending
``bandmatrix_sparse``              Spectrum.bandmatrix(H, ks; format=:sparse, num_bands)
                                   (Spectrum/bands.jl:164-227)
``eigen_sparse_matrix(M, ...)``    Eigen.eigen_sparse on one sparse matrix
``eigmin_sparse``/``eigmax_sparse`` (eigen_sparse_krylovkit.jl:91-92)
=================================  ==========================================================

The device holds one CSR for all keys (entry = column, key, value), so H(k) is never
materialised: the solver's sparse products apply it to blocks of vectors of many k-points at
once, each k-point only bringing its phase table (``lqm_eigen_partial``).
"""
from __future__ import annotations

import ctypes as _c
import warnings

import numpy as np
import scipy.sparse as sp

from . import _lib
from .hops import Hops, _key, points

_WHICH = {"LM": 0, "SR": 1, "LR": 2}


class SparseHops:
    """R (tuple of ints) -> sparse N x N complex block; H(k) = sum_R t_R e^{2 pi i k.R}."""

    def __init__(self, data=None):
        self.data = {}
        if data is None:
            return
        items = data.items() if isinstance(data, dict) else data
        for R, t in items:
            self[R] = t

    @classmethod
    def from_hops(cls, h: Hops) -> "SparseHops":
        return cls({R: t for R, t in h.items()})

    def __setitem__(self, R, t) -> None:
        t = sp.csr_matrix(t, dtype=np.complex128)
        if t.shape[0] != t.shape[1]:
            raise ValueError("hopping blocks must be square matrices")
        t.eliminate_zeros()
        self.data[_key(R)] = t

    def __getitem__(self, R):
        return self.data[_key(R)]

    def __contains__(self, R) -> bool:
        return _key(R) in self.data

    def keys(self):
        return self.data.keys()

    def items(self):
        return self.data.items()

    @property
    def N(self) -> int:
        return next(iter(self.data.values())).shape[0]

    @property
    def d(self) -> int:
        return len(next(iter(self.data.keys())))

    @property
    def nnz(self) -> int:
        return sum(t.nnz for t in self.data.values())

    def __call__(self, k):
        """H(k) as a scipy CSR matrix."""
        k = np.asarray(k, dtype=float).ravel()
        out = sp.csr_matrix((self.N, self.N), dtype=np.complex128)
        for R, t in self.data.items():
            out = out + t * np.exp(2j * np.pi * float(np.dot(k, R)))
        return out

    def ishermitian(self, tol: float = 0.0) -> bool:
        for R, t in self.data.items():
            mR = tuple(-x for x in R)
            if mR not in self.data:
                return False
            if abs(t - self.data[mR].conj().T).max() > tol:
                return False
        return True

    def csr(self):
        """(R [d, nR] int32, rowptr [N+1] int64, col int32, key int32, val complex128): the
        entries of every key merged row by row (key order, then column, within a row)."""
        keys = list(self.data.keys())
        # column r = key r (column-major, the components of a key contiguous)
        R = np.asfortranarray(np.array(keys, dtype=np.int32).reshape(len(keys), self.d).T)
        rows, cols, kk, vals = [], [], [], []
        for r, key in enumerate(keys):
            c = self.data[key].tocoo()
            rows.append(c.row.astype(np.int64))
            cols.append(c.col.astype(np.int32))
            kk.append(np.full(c.nnz, r, dtype=np.int32))
            vals.append(c.data.astype(np.complex128))
        row = np.concatenate(rows) if rows else np.zeros(0, np.int64)
        col = np.concatenate(cols) if cols else np.zeros(0, np.int32)
        key = np.concatenate(kk) if kk else np.zeros(0, np.int32)
        val = np.concatenate(vals) if vals else np.zeros(0, np.complex128)
        order = np.lexsort((col, key, row))
        row, col, key, val = row[order], col[order], key[order], val[order]
        rowptr = np.zeros(self.N + 1, dtype=np.int64)
        np.add.at(rowptr, row + 1, 1)
        rowptr = np.cumsum(rowptr)
        return (R, rowptr, np.ascontiguousarray(col), np.ascontiguousarray(key), np.ascontiguousarray(val))


class SparseModel:
    """A SparseHops uploaded once and kept resident on the device (lqm_sparse_model)."""

    def __init__(self, H, executor=None):
        from .api import to_executor

        if isinstance(H, Hops):
            H = SparseHops.from_hops(H)
        self.exec = executor if executor is not None and not isinstance(executor, str) else to_executor(executor or "gpu")
        R, rowptr, col, key, val = H.csr()
        self.N, self.d, self.nR = H.N, H.d, R.shape[1]
        h = _c.c_void_p()
        vv = np.ascontiguousarray(val.view(np.float64))
        _lib.call("lqm_sparse_model_create", self.exec.handle, self.N, self.d, self.nR, R.ctypes.data_as(_c.c_void_p),
                  int(col.size), rowptr.ctypes.data_as(_c.c_void_p), col.ctypes.data_as(_c.c_void_p),
                  key.ctypes.data_as(_c.c_void_p), vv.ctypes.data_as(_c.c_void_p), _c.byref(h))
        self.handle = h

    @property
    def bound(self) -> float:
        b = _c.c_double()
        _lib.call("lqm_sparse_model_bound", self.handle, _c.byref(b))
        return b.value

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            _lib.lib().lqm_sparse_model_destroy(self.handle)
            self.handle = _c.c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def eigen_sparse(H, ks, num_bands: int, sigma: float = 1e-8, which: str = "LM", tol: float = 1e-12,
                 maxiter: int = 100, vectors: bool = True, executor=None, return_iters: bool = False):
    """(bands [num_bands, Nk], vecs [N, num_bands, Nk] or None): per k-point the eigenpairs the
    reference's eigen_sparse returns, eigenvalues ascending.  A k-point that does not converge
    within maxiter iterations returns its last Ritz pairs with a warning (as the reference)."""
    if which not in _WHICH:
        raise ValueError(f"unsupported which={which!r} (use 'LM', 'SR' or 'LR')")
    m = H if isinstance(H, SparseModel) else SparseModel(H, executor)
    k = np.asfortranarray(points(ks), dtype=np.float64)
    if k.shape[0] != m.d:
        raise ValueError(f"k-points have dimension {k.shape[0]}, the model has {m.d}")
    Nk = k.shape[1]
    bands = np.zeros((num_bands, Nk), order="F")
    vecs = np.zeros((m.N, num_bands, Nk), dtype=np.complex128, order="F") if vectors else None
    iters = np.zeros(Nk, dtype=np.int32)
    _lib.call("lqm_eigen_partial", m.handle, Nk, k.ctypes.data_as(_c.c_void_p), int(num_bands), _WHICH[which],
              float(sigma), float(tol), int(maxiter), bands.ctypes.data_as(_c.c_void_p),
              vecs.ctypes.data_as(_c.c_void_p) if vectors else None, iters.ctypes.data_as(_c.c_void_p))
    if np.any(iters < 0):
        warnings.warn(f"eigen_sparse: {int(np.sum(iters < 0))} of {Nk} k-points did not converge within "
                      f"{maxiter} iterations (which={which}, sigma={sigma})")
    out = (bands, vecs)
    return out + (iters,) if return_iters else out


def bandmatrix_sparse(H, ks, num_bands: int, sigma: float = 1e-8, which: str = "LM", **kw) -> np.ndarray:
    """Bands only: [num_bands, Nk]."""
    return eigen_sparse(H, ks, num_bands, sigma=sigma, which=which, vectors=False, **kw)[0]


def eigen_sparse_matrix(M, num_bands: int, sigma: float = 1e-8, which: str = "LM", **kw):
    """eigen_sparse on one Hermitian matrix: (eigvals [num_bands], eigvecs [N, num_bands])."""
    M = sp.csr_matrix(M, dtype=np.complex128)
    b, v = eigen_sparse(SparseHops({(0,): M}), np.zeros((1, 1)), num_bands, sigma=sigma, which=which, **kw)
    return b[:, 0], v[:, :, 0]


def eigmin_sparse(M, **kw) -> float:
    return float(eigen_sparse_matrix(M, 1, which="SR", **kw)[0][0])


def eigmax_sparse(M, **kw) -> float:
    return float(eigen_sparse_matrix(M, 1, which="LR", **kw)[0][0])
