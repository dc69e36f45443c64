"""FP64 CPU oracle of the partial spectrum -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` may import this module, as the checker; the product never calls it.
Restates what Eigen.eigen_sparse returns (Eigen/eigen_sparse_krylovkit.jl:44-88) from the
full spectrum of H(k) (LAPACK zheevr through the C oracle): which = "LM" the num_bands
eigenvalues nearest sigma (the set sort(e, by = |e - sigma|)[1:num_bands] of the reference's
test, test/test_eigen.jl:24), "SR" the smallest, "LR" the largest -- every selection returned
ascending (eigen_sparse_krylovkit.jl:82-87), with the eigenvectors of the selected values.
"""
from __future__ import annotations

import numpy as np

from . import oracle as _o


def select(ev, num_bands, sigma=1e-8, which="LM"):
    """Indices of the selected eigenvalues of an ascending vector, ascending."""
    ev = np.asarray(ev)
    if which == "LM":
        idx = np.argsort(np.abs(ev - sigma), kind="stable")[:num_bands]
    elif which == "SR":
        idx = np.arange(num_bands)
    elif which == "LR":
        idx = np.arange(ev.size - num_bands, ev.size)
    else:
        raise ValueError(which)
    return np.sort(idx)


def eigen_partial(R, T, ks, num_bands, sigma=1e-8, which="LM", vectors=False, nthreads=None):
    out = _o.bands(R, T, np.asfortranarray(ks), vectors=vectors, nthreads=nthreads)
    b, V = out if vectors else (out, None)
    Nk = b.shape[1]
    bands = np.zeros((num_bands, Nk))
    vecs = np.zeros((b.shape[0], num_bands, Nk), dtype=np.complex128) if vectors else None
    for k in range(Nk):
        idx = select(b[:, k], num_bands, sigma, which)
        bands[:, k] = b[idx, k]
        if vectors:
            vecs[:, :, k] = V[:, idx, k]
    return (bands, vecs) if vectors else bands
