"""FP64 CPU oracle for the k-space band path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker / the timed CPU
baseline.  The product library never calls it.

The per-k loops (Bloch sum, LAPACK ``zheevr``, DOS, density matrix) live in
``lqm_oracle.c``; this module binds them with ctypes and adds the scalar
pieces of the reference that are cheap in numpy:

* ``chemicalpotential`` -- Spectrum/chemicalpotential.jl:6-46 (T=0 order
  statistic; T>0 bracketing bisection on getfilling, the method Roots.jl's
  ``find_zero(f, (a, b))`` uses for a bracket).
* ``getfilling``        -- Spectrum/chemicalpotential.jl:74-78.
* ``fermidirac``        -- Utils/fermidirac.jl:1-5.

Layouts follow Julia (column-major); numpy arrays are passed as Fortran-ordered
buffers: hopping stack ``T`` of shape (N, N, nR), keys ``R`` (d, nR) int32,
k-points ``ks`` (d, Nk), bands (N, Nk), rho (N, N, nRr).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liblqm_oracle.so")
_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _capsule_ptr(capsule) -> int:
    get_name = ctypes.pythonapi.PyCapsule_GetName
    get_name.restype = ctypes.c_char_p
    get_name.argtypes = [ctypes.py_object]
    get_ptr = ctypes.pythonapi.PyCapsule_GetPointer
    get_ptr.restype = ctypes.c_void_p
    get_ptr.argtypes = [ctypes.py_object, ctypes.c_char_p]
    return get_ptr(capsule, get_name(capsule))


def lib():
    """Load the oracle library (building it if absent) and wire LAPACK/BLAS."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = ctypes.CDLL(_LIB_PATH)
    import scipy.linalg.cython_blas as cb
    import scipy.linalg.cython_lapack as cl

    L.lqo_set_lapack.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.lqo_set_lapack(_capsule_ptr(cl.__pyx_capi__["zheevr"]), _capsule_ptr(cb.__pyx_capi__["zgemm"]))
    P = ctypes.c_void_p
    i = ctypes.c_int
    dbl = ctypes.c_double
    L.lqo_fouriersum.argtypes = [i, i, i, P, P, P, P]
    L.lqo_eigh.argtypes = [i, P, P, P]
    L.lqo_eigh.restype = i
    L.lqo_bands.argtypes = [i, i, i, P, P, i, P, P, P, i]
    L.lqo_bands.restype = i
    L.lqo_dos.argtypes = [i, i, i, P, P, i, P, P, i, P, dbl, P, i]
    L.lqo_dos.restype = i
    L.lqo_densitymatrix.argtypes = [i, i, i, P, P, i, P, P, dbl, dbl, i, P, P, P, i]
    L.lqo_densitymatrix.restype = i
    L.lqo_fermidirac.argtypes = [dbl, dbl, dbl]
    L.lqo_fermidirac.restype = dbl
    L.lqo_max_threads.restype = i
    _lib = L
    return L


class _OneBlasThread:
    """Pin OpenBLAS to one thread per caller thread (Parallel.configure_blas!)."""

    def __enter__(self):
        lib()  # load scipy's OpenBLAS first so that the limit below applies to it
        try:
            from threadpoolctl import threadpool_limits

            self._ctx = threadpool_limits(limits=1, user_api="blas")
        except Exception:  # pragma: no cover - threadpoolctl is part of the image
            self._ctx = None
        return self

    def __exit__(self, *exc):
        if self._ctx is not None:
            self._ctx.unregister() if hasattr(self._ctx, "unregister") else self._ctx.__exit__(*exc)
        return False


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _pack(R, T):
    R = np.asfortranarray(np.asarray(R, dtype=np.int32))
    T = np.asfortranarray(np.asarray(T, dtype=np.complex128))
    assert T.ndim == 3 and T.shape[0] == T.shape[1] and T.shape[2] == R.shape[1]
    return R, T


def default_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def fouriersum(R, T, k) -> np.ndarray:
    """H(k) = sum_R T_R exp(2 pi i k.R)   (TightBinding/types.jl:18-25)."""
    R, T = _pack(R, T)
    k = np.ascontiguousarray(np.asarray(k, dtype=np.float64))
    N = T.shape[0]
    out = np.zeros((N, N), dtype=np.complex128, order="F")
    lib().lqo_fouriersum(N, R.shape[0], R.shape[1], _ptr(R), _ptr(T), _ptr(k), _ptr(out))
    return out


def eigh(A, vectors=True):
    """LAPACK zheevr on Hermitian(A) (upper triangle), as Eigen.jl:14-39."""
    A = np.asfortranarray(np.asarray(A, dtype=np.complex128))
    N = A.shape[0]
    w = np.zeros(N)
    V = np.zeros((N, N), dtype=np.complex128, order="F") if vectors else None
    with _OneBlasThread():
        st = lib().lqo_eigh(N, _ptr(A), _ptr(w), _ptr(V) if vectors else None)
    if st:
        raise RuntimeError(f"zheevr failed (info={st})")
    return (w, V) if vectors else w


def bands(R, T, ks, vectors=False, nthreads=None):
    """bandmatrix(H, ks) without projectors: (N, Nk) ascending eigenvalues."""
    R, T = _pack(R, T)
    ks = np.asfortranarray(np.asarray(ks, dtype=np.float64))
    N, Nk = T.shape[0], ks.shape[1]
    ev = np.zeros((N, Nk), order="F")
    V = np.zeros((N, N, Nk), dtype=np.complex128, order="F") if vectors else None
    with _OneBlasThread():
        st = lib().lqo_bands(N, R.shape[0], R.shape[1], _ptr(R), _ptr(T), Nk, _ptr(ks), _ptr(ev),
                             _ptr(V) if vectors else None, nthreads or default_threads())
    if st:
        raise RuntimeError(f"oracle bands failed ({st})")
    return (ev, V) if vectors else ev


def dos(R, T, ks, omegas, gamma, kweights=None, nthreads=None):
    """dos_compute! (Spectrum/dos.jl:202-232)."""
    R, T = _pack(R, T)
    ks = np.asfortranarray(np.asarray(ks, dtype=np.float64))
    om = np.ascontiguousarray(np.asarray(omegas, dtype=np.float64))
    N, Nk = T.shape[0], ks.shape[1]
    out = np.zeros(len(om))
    kw = None if kweights is None else np.ascontiguousarray(np.asarray(kweights, dtype=np.float64))
    with _OneBlasThread():
        st = lib().lqo_dos(N, R.shape[0], R.shape[1], _ptr(R), _ptr(T), Nk, _ptr(ks),
                           _ptr(kw) if kw is not None else None, len(om), _ptr(om), float(gamma),
                           _ptr(out), nthreads or default_threads())
    if st:
        raise RuntimeError(f"oracle dos failed ({st})")
    return out


def densitymatrix(R, T, ks, Rr, mu=0.0, Temp=0.01, kweights=None, nthreads=None):
    """getdensitymatrix! (Operators/densitymatrix.jl:17-145): returns (rho, E_band)."""
    R, T = _pack(R, T)
    Rr = np.asfortranarray(np.asarray(Rr, dtype=np.int32))
    ks = np.asfortranarray(np.asarray(ks, dtype=np.float64))
    N, Nk = T.shape[0], ks.shape[1]
    rho = np.zeros((N, N, Rr.shape[1]), dtype=np.complex128, order="F")
    eb = np.zeros(1)
    kw = None if kweights is None else np.ascontiguousarray(np.asarray(kweights, dtype=np.float64))
    with _OneBlasThread():
        st = lib().lqo_densitymatrix(N, R.shape[0], R.shape[1], _ptr(R), _ptr(T), Nk, _ptr(ks),
                                     _ptr(kw) if kw is not None else None, float(mu), float(Temp),
                                     Rr.shape[1], _ptr(Rr), _ptr(rho), _ptr(eb),
                                     nthreads or default_threads())
    if st:
        raise RuntimeError(f"oracle densitymatrix failed ({st})")
    return rho, float(eb[0])


def fermidirac(e, mu=0.0, T=0.0):
    """Utils/fermidirac.jl:1-5 (heaviside with 1/2 at zero for T=0)."""
    x = np.asarray(e, dtype=np.float64) - mu
    if T > 0:
        with np.errstate(over="ignore"):
            return 1.0 / (np.exp(x / T) + 1.0)
    return np.where(x < 0, 1.0, np.where(x > 0, 0.0, 0.5))


def getfilling(bands, mu, T=0.0, kweights=None):
    """Spectrum/chemicalpotential.jl:74-78."""
    f = fermidirac(bands, mu, T)
    if kweights is None:
        return float(np.mean(f))
    return float(np.sum(np.asarray(kweights)[None, :] * f) / f.shape[0])


def _middle(a, b):
    """Roots.jl's bracketing midpoint for Float64: 0 across a sign change, else the
    midpoint of the IEEE bit patterns of |a| and |b| with the sign of a + b, so the
    bisection reaches adjacent doubles in at most 64 steps."""
    if (a < 0) != (b < 0) and a != 0 and b != 0:
        return 0.0
    ia = int(np.array(abs(a), dtype=np.float64).view(np.uint64))
    ib = int(np.array(abs(b), dtype=np.float64).view(np.uint64))
    mid = float(np.array((ia + ib) >> 1, dtype=np.uint64).view(np.float64))
    return -mid if (a + b) < 0 else mid


def _bisect_doubles(fn, a, b):
    """Bracketing bisection with Roots.jl's midpoint down to adjacent doubles; a midpoint
    with fn == 0 is returned at once, else the end with the smaller |fn|."""
    fa, fb = fn(a), fn(b)
    if fa == 0:
        return a
    if fb == 0:
        return b
    for _ in range(200):
        m = _middle(a, b)
        if not (a < m < b):
            break
        fm = fn(m)
        if fm == 0:
            return m
        if (fm < 0) == (fa < 0):
            a, fa = m, fm
        else:
            b, fb = m, fm
    return a if abs(fa) <= abs(fb) else b


def chemicalpotential(bands, filling, T=0.0, kweights=None):
    """chemicalpotential(bands, kgrid, filling; T) (Spectrum/chemicalpotential.jl:6-46)."""
    bands = np.asarray(bands, dtype=np.float64)
    if T == 0:
        en = np.sort(bands.ravel(order="F"))
        if filling >= 1.0:
            return float(en[-1])
        if filling <= 0.0:
            return float(en[0] - 1.0)
        i = 1 + int(np.floor(filling * (len(en) - 1)))  # 1-based, chemicalpotential.jl:20
        return float((en[i - 1] + en[i]) / 2)
    lo, hi = float(bands.min()), float(bands.max())
    return _bisect_doubles(lambda m: getfilling(bands, m, T, kweights) - filling, lo, hi)
