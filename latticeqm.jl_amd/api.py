"""Host API of the GPU band path, mirroring the reference's entry points.

Each function has the name, argument meaning and error behaviour of its
LatticeQM.jl counterpart, with ``multimode=:gpu`` as the only executor:

=====================  =================================================
this module            reference
=====================  =================================================
``to_executor``        Parallel.to_executor (Parallel/Parallel.jl:85-109)
``fouriersum``         TightBinding.fouriersum! (TightBinding/types.jl:18-25)
``bandmatrix``         Spectrum.bandmatrix (Spectrum/bands.jl:164-204)
``getbands``           Spectrum.getbands (Spectrum/bands.jl:264-267)
``getdos``             Spectrum.getdos (Spectrum/dos.jl:127-232)
``chemicalpotential``  Spectrum.chemicalpotential (Spectrum/chemicalpotential.jl:6-65)
``getfilling``         Spectrum.getfilling / filling (chemicalpotential.jl:74-94)
``getdensitymatrix``   Operators.getdensitymatrix! (Operators/densitymatrix.jl:17-49)
``setfilling``         Operators.setfilling! (Operators/chemicalpotential.jl:20-46)
=====================  =================================================

All numerics run in liblqmgpu.so on the GPU; this module only packs arrays and
maps status codes to exceptions.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from .hops import Hops, Mesh, points, regulargrid

_c = ctypes


def _vp(a: Optional[np.ndarray]):
    return None if a is None else _c.c_void_p(a.ctypes.data)


class GPUExec:
    """The executor multimode=:gpu resolves to: one device context + stream."""

    def __init__(self, device: int = 0, precision: int = _lib.LQM_PREC_FP64):
        self.device = device
        self.precision = precision
        h = _c.c_void_p()
        _lib.call("lqm_context_create", device, precision, _c.byref(h))
        self.handle = h

    def synchronize(self) -> None:
        _lib.call("lqm_context_synchronize", self.handle)

    def set_variant(self, kind: int, variant: int) -> None:
        """Force kernel variant ``variant`` (-1: tuned/default) for kind 0 (tridiagonalization)
        or kind 1 (rho_k construction)."""
        _lib.call("lqm_context_set_variant", self.handle, int(kind), int(variant))

    def get_variant(self, kind: int, N: int) -> int:
        v = _c.c_int()
        _lib.call("lqm_context_get_variant", self.handle, int(kind), int(N), _c.byref(v))
        return v.value

    def set_refinement(self, max_sweeps: int) -> None:
        """Largest number of Ogita-Aishima sweeps of the mixed-precision path."""
        _lib.call("lqm_context_set_refinement", self.handle, int(max_sweeps))

    def set_refinement_threshold(self, thr: float) -> None:
        """Squared-correction level below which a k-point stops refining (default 1e-8;
        <= 0: two sweeps for every k-point)."""
        _lib.call("lqm_context_set_refinement_threshold", self.handle, float(thr))

    def refinement_report(self) -> dict:
        """Diagnostics of the last mixed-precision call."""
        out = (_c.c_double * 6)()
        _lib.call("lqm_context_refinement_report", self.handle, out)
        return {"sweeps": int(out[0]), "max_orth_residual": out[1], "max_cluster_delta": out[2],
                "kpoints_near_degenerate": int(out[3]), "kpoint_sweeps": int(out[4]),
                "kpoints_rho_fused": int(out[5])}

    def save_tuning(self, path) -> None:
        """Write the autotuned variants (lqm_context_save_tuning)."""
        _lib.call("lqm_context_save_tuning", self.handle, str(path).encode())

    def load_tuning(self, path) -> None:
        """Merge autotuned variants from a file written on this architecture."""
        _lib.call("lqm_context_load_tuning", self.handle, str(path).encode())

    def set_graphs(self, enable: bool) -> None:
        """Replay SCF iterations from captured HIP graphs where eligible (default on)."""
        _lib.call("lqm_context_set_graphs", self.handle, int(bool(enable)))

    def last_precision(self) -> str | None:
        """Precision the last pass actually ran in ("fp64" / "mixed"; None before any pass).
        A mixed executor runs FP64 for eigenvector output and for the large-N layout."""
        v = _c.c_int()
        _lib.call("lqm_context_last_precision", self.handle, _c.byref(v))
        return {0: "fp64", 1: "mixed"}.get(v.value)

    def set_stream(self, stream_ptr: int | None) -> None:
        _lib.call("lqm_context_set_stream", self.handle, _c.c_void_p(stream_ptr or 0))

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            _lib.lib().lqm_context_destroy(self.handle)
            self.handle = _c.c_void_p()

    def __del__(self):  # pragma: no cover - interpreter shutdown order varies
        try:
            self.close()
        except Exception:
            pass


_executors: dict = {}


_PRECISIONS = {"fp64": _lib.LQM_PREC_FP64, "mixed": _lib.LQM_PREC_FP32_REFINED}


def to_executor(multimode="gpu", device: int = 0, precision: str = "fp64") -> GPUExec:
    """Resolve ``multimode`` (``:gpu``/``"gpu"`` or a GPUExec) to a shared executor.

    precision "fp64" is the plain path; "mixed" runs the eigensolve in FP32 and
    refines the eigenpairs to FP64 accuracy (Ogita-Aishima)."""
    if isinstance(multimode, GPUExec):
        return multimode
    mode = str(multimode).lstrip(":")
    if mode != "gpu":
        raise ValueError(f"this package provides multimode=:gpu only, got {multimode!r}")
    if precision not in _PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}")
    key = (device, precision)
    ex = _executors.get(key)
    if ex is None:
        ex = _executors[key] = GPUExec(device, _PRECISIONS[precision])
    return ex


class GPUModel:
    """A Hops operator uploaded once and kept resident in device memory."""

    def __init__(self, H: Hops, executor: GPUExec | None = None):
        self.exec = executor or to_executor()
        R, T = H.pack()
        self.R, self.N, self.d, self.nR = R, H.N, H.d, R.shape[1]
        self.keys = list(H.keys())
        h = _c.c_void_p()
        _lib.call("lqm_model_create", self.exec.handle, self.N, self.d, self.nR, _vp(R), _vp(T),
                  _c.byref(h))
        self.handle = h

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            _lib.lib().lqm_model_destroy(self.handle)
            self.handle = _c.c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def autotune(model: GPUModel, ks_dev_ptr: int, Nk: int) -> dict:
    """Time every kernel variant on Nk device-resident k-points and keep the fastest."""
    _lib.call("lqm_autotune", model.handle, int(Nk), _c.c_void_p(ks_dev_ptr))
    return {kind: model.exec.get_variant(kind, model.N) for kind in (0, 1, 3, 5, 6)}


def _model(H, executor) -> GPUModel:
    if isinstance(H, GPUModel):
        return H
    if not isinstance(H, Hops):
        raise TypeError("H must be a Hops or a GPUModel")
    return GPUModel(H, to_executor(executor or "gpu"))


def _kpts(ks, d: int) -> np.ndarray:
    k = np.asfortranarray(points(ks), dtype=np.float64)
    if k.shape[0] != d:
        raise ValueError(f"k-points have dimension {k.shape[0]}, the model has {d}")
    return k


def _weights(ks, kweights, Nk: int | None = None):
    """Per-k weights as a contiguous float64 array (None: uniform 1/Nk).  The library copies
    exactly Nk doubles, so a weight vector of any other length is an error here."""
    if kweights is not None:
        w = np.ascontiguousarray(kweights, dtype=np.float64)
    elif isinstance(ks, Mesh):
        w = np.ascontiguousarray(ks.weights(), dtype=np.float64)
    else:
        return None
    if Nk is not None and w.shape != (Nk,):
        raise ValueError(f"kweights must have shape ({Nk},), got {w.shape}")
    return w


def fouriersum(H, ks, executor=None) -> np.ndarray:
    """H(k) for every column of ks: complex array (N, N, Nk)."""
    m = _model(H, executor)
    k = _kpts(ks, m.d)
    Nk = k.shape[1]
    out = np.zeros((m.N, m.N, Nk), dtype=np.complex128, order="F")
    _lib.call("lqm_fouriersum", m.handle, Nk, _vp(k), _vp(out))
    return out


def _projector_models(projectors, m: GPUModel):
    """handleprojector (bands.jl:62-66): a matrix, a Hops operator or a list of them."""
    if len(projectors) == 1 and isinstance(projectors[0], (list, tuple)):
        projectors = tuple(projectors[0])
    out = []
    for P in projectors:
        if isinstance(P, Hops):
            out.append(GPUModel(P, m.exec))
        elif isinstance(P, np.ndarray) and P.ndim == 2:
            out.append(GPUModel(Hops({(0,) * m.d: P}), m.exec))
        else:
            raise NotImplementedError("projectors must be matrices or Hops operators")
    return out


def bandmatrix(H, ks, *projectors, executor=None, multimode="gpu", format="dense", num_bands=None,
               sigma=1e-8, **kw):
    """Eigenvalues of Hermitian(H(k)) per k: returns (bands (N, Nk), obs (N, Nk, nP)) with
    obs[n, j, p] = real(psi^H O_p(k_j) psi) for the eigenvector of bands[n, j].
    format="sparse" with num_bands: the num_bands eigenvalues nearest sigma per k (the
    partial-spectrum path), bands (num_bands, Nk)."""
    if format == "sparse":
        from .sparse import eigen_sparse

        if num_bands is None:
            raise ValueError("format='sparse' needs num_bands")
        if projectors:
            raise NotImplementedError("projectors with format='sparse'")
        b, _ = eigen_sparse(H, ks, num_bands, sigma=sigma, vectors=False, executor=executor, **kw)
        return b, np.zeros((num_bands, b.shape[1], 0))
    m = _model(H, executor or multimode)
    k = _kpts(ks, m.d)
    Nk = k.shape[1]
    bands = np.zeros((m.N, Nk), order="F")
    if not projectors:
        _lib.call("lqm_bandmatrix", m.handle, Nk, _vp(k), _vp(bands), None)
        return bands, np.zeros((m.N, Nk, 0))
    pm = _projector_models(projectors, m)
    obs = np.zeros((m.N, Nk, len(pm)), order="F")
    handles = (_c.c_void_p * len(pm))(*[p.handle.value for p in pm])
    _lib.call("lqm_bandmatrix_expval", m.handle, Nk, _vp(k), len(pm), handles, _vp(bands), _vp(obs))
    return bands, obs


def eigen(H, ks, executor=None, format="dense", num_bands=None, sigma=1e-8, **kw):
    """Eigenvalues and eigenvectors per k (the geteigen! branch, bands.jl:222-224):
    (bands (N, Nk), vecs (N, N, Nk)) with vecs[:, n, j] the eigenvector of bands[n, j].
    format="sparse" with num_bands: the num_bands eigenpairs nearest sigma per k."""
    if format == "sparse":
        from .sparse import eigen_sparse

        if num_bands is None:
            raise ValueError("format='sparse' needs num_bands")
        return eigen_sparse(H, ks, num_bands, sigma=sigma, executor=executor, **kw)
    m = _model(H, executor)
    k = _kpts(ks, m.d)
    Nk = k.shape[1]
    bands = np.zeros((m.N, Nk), order="F")
    vecs = np.zeros((m.N, m.N, Nk), dtype=np.complex128, order="F")
    _lib.call("lqm_bandmatrix", m.handle, Nk, _vp(k), _vp(bands), _vp(vecs))
    return bands, vecs


@dataclass
class BandData:
    """Spectrum.BandData (Spectrum/types.jl:23-29)."""
    bands: np.ndarray
    obs: np.ndarray
    path: object = field(default=None)


def getbands(H, ks, *projectors, **kw) -> BandData:
    bands, obs = bandmatrix(H, ks, *projectors, **kw)
    return BandData(bands, obs, ks)


def getdos(h, *args, klin: int | None = None, Gamma: float | None = None, kweights=None,
           executor=None, **kw):
    """DOS with a Lorentzian of width Gamma and no 1/pi (Spectrum/dos.jl:127-232).

    getdos(h, emin, emax, num=500; klin, Gamma) -> (omegas, DOS)
    getdos(h, omegas; klin, Gamma)              -> DOS
    getdos(h, omegas, ks; Gamma, kweights)       -> DOS
    """
    if Gamma is None:
        Gamma = kw.pop("Γ", None)
    if Gamma is None:
        raise ValueError("keyword Gamma (Γ) is required")
    if len(args) >= 2 and np.isscalar(args[0]) and np.isscalar(args[1]):
        emin, emax = float(args[0]), float(args[1])
        num = int(args[2]) if len(args) > 2 else 500
        om = np.linspace(emin, emax, num)
        return om, getdos(h, om, klin=klin, Gamma=Gamma, executor=executor)
    om = np.ascontiguousarray(args[0], dtype=np.float64)
    if len(args) > 1:
        ks = args[1]
    else:
        if klin is None:
            raise ValueError("keyword klin is required")
        ks = regulargrid(nk=klin ** 2)
    m = _model(h, executor)
    k = _kpts(ks, m.d)
    w = _weights(ks, kweights, k.shape[1])
    out = np.zeros(len(om))
    _lib.call("lqm_dos", m.handle, k.shape[1], _vp(k), _vp(w), len(om), _vp(om), float(Gamma), _vp(out))
    return out


def ldos(H, ks, frequencies, Gamma: float | None = None, kweights=None, executor=None, **kw) -> np.ndarray:
    """Site-resolved local DOS (Spectrum/dos.jl:57-105): k-averaged, averaged over the
    frequency window, with the spectral function's 1/pi, so that at one frequency the sum over
    sites is getdos / pi.  ``Gamma`` (or ``Γ``) defaults to 0.1 as in the reference."""
    if Gamma is None:
        Gamma = kw.pop("Γ", 0.1)
    om = np.atleast_1d(np.ascontiguousarray(frequencies, dtype=np.float64))
    if om.size == 0:
        raise ValueError("at least one frequency")
    m = _model(H, executor)
    k = _kpts(ks, m.d)
    w = _weights(ks, kweights, k.shape[1])
    out = np.zeros(m.N)
    _lib.call("lqm_ldos", m.handle, k.shape[1], _vp(k), _vp(w), om.size, _vp(om), float(Gamma), _vp(out))
    return out


def density(H, ks, mu: float = 0.0, kweights=None, executor=None) -> np.ndarray:
    """Site densities of the states with e <= mu, k-averaged (Spectrum/dos.jl:20-52)."""
    m = _model(H, executor)
    k = _kpts(ks, m.d)
    w = _weights(ks, kweights, k.shape[1])
    out = np.zeros(m.N)
    _lib.call("lqm_density", m.handle, k.shape[1], _vp(k), _vp(w), float(mu), _vp(out))
    return out


def chemicalpotential(H, ks, filling: float, T: float = 0.0, executor=None, kweights=None) -> float:
    """mu such that the filling of the bands on ks equals ``filling``."""
    m = _model(H, executor)
    k = _kpts(ks, m.d)
    w = _weights(ks, kweights, k.shape[1])
    mu = _c.c_double()
    _lib.call("lqm_chemicalpotential", m.handle, k.shape[1], _vp(k), _vp(w), float(filling), float(T),
              _c.byref(mu))
    return mu.value


def getdensitymatrix(rhos: Hops, H, ks, mu: float = 0.0, T: float = 0.01, kweights=None,
                     executor=None) -> float:
    """Overwrite every block of ``rhos`` with rho(R) and return E_band.

    rho(R) = sum_k w_k e^{-2 pi i k.R} U_k f(e - mu) U_k^H with w_k = 1/Nk (or the
    Mesh / explicit weights); only the keys already present in ``rhos`` are filled.
    """
    m = _model(H, executor)
    k = _kpts(ks, m.d)
    w = _weights(ks, kweights, k.shape[1])
    keys = list(rhos.keys())
    if keys and len(keys[0]) != m.d:
        raise ValueError("rho keys and model keys differ in dimension")
    for key in keys:
        if rhos[key].shape != (m.N, m.N):
            raise ValueError("rho blocks must be N x N")
    nRr = len(keys)
    Rr = np.asfortranarray(np.array(keys, dtype=np.int32).T.reshape(m.d, nRr)) if nRr else np.zeros((m.d, 1), np.int32)
    rho = np.zeros((m.N, m.N, max(nRr, 1)), dtype=np.complex128, order="F")
    eb = _c.c_double()
    _lib.call("lqm_densitymatrix", m.handle, k.shape[1], _vp(k), _vp(w), float(mu), float(T), nRr,
              _vp(Rr), _vp(rho), _c.byref(eb))
    for r, key in enumerate(keys):
        rhos.data[key] = np.array(rho[:, :, r])
    return eb.value


def addchemicalpotential(H: Hops, mu: float) -> Hops:
    """Add mu to the onsite diagonal (Operators/chemicalpotential.jl:48-58)."""
    zk = H.zerokey()
    if zk not in H:
        H[zk] = np.zeros((H.N, H.N), dtype=np.complex128)
    blk = H.data[zk]
    blk[np.diag_indices(H.N)] += mu
    return H


def setfilling(H: Hops, filling: float, nk: int = 100, T: float = 0.0, executor=None) -> float:
    """Shift H so that its filling on regulargrid(nk) is ``filling``; returns mu."""
    mu = chemicalpotential(H, regulargrid(nk=nk, dim=H.d), filling, T=T, executor=executor)
    addchemicalpotential(H, -mu)
    return mu


# ---- spectral adjoints: grand potential and band-loss gradients --------------------
def grandpotential(H, ks, mu: float = 0.0, T: float = 0.05, kweights=None, executor=None,
                   gradient: bool = False):
    """Omega = sum_k w_k sum_n -T ln(1 + exp(-(e_nk - mu)/T)) (min(e - mu, 0) at T = 0).

    With ``gradient=True`` returns ``(Omega, dOmega)``, dOmega a Hops on the keys of H with
    dOmega[R][i, j] = dOmega/d(t_R)_ij = rho(-R)[j, i] (Hellmann-Feynman)."""
    m = _model(H, executor)
    k = _kpts(ks, m.d)
    w = _weights(ks, kweights, k.shape[1])
    om = _c.c_double()
    g = np.zeros((m.N, m.N, m.nR), dtype=np.complex128, order="F") if gradient else None
    _lib.call("lqm_grandpotential", m.handle, k.shape[1], _vp(k), _vp(w), float(mu), float(T),
              _c.byref(om), _vp(g))
    if not gradient:
        return om.value
    return om.value, Hops({key: np.array(g[:, :, r]) for r, key in enumerate(m.keys)})


def grandpotential_rrule(H: Hops, ks, mu: float = 0.0, T: float = 0.05, kweights=None, executor=None):
    """(Omega, pullback): pullback(dOmega_bar) returns the cotangent of every block t_R,
    dOmega_bar * conj(dOmega/dt_R) = dOmega_bar * rho(R) -- the gradient of the real Omega
    in the real inner product <X, D> = Re sum conj(X) D, the convention of ChainRulesCore
    for complex inputs."""
    om, g = grandpotential(H, ks, mu, T, kweights, executor, gradient=True)

    def pullback(dbar: float = 1.0) -> Hops:
        return Hops({R: dbar * np.conj(b) for R, b in g.items()})

    return om, pullback


def densitymatrix_occ(rhos: Hops, H, ks, occ, kweights=None, executor=None) -> Hops:
    """Overwrite rhos with sum_k w_k e^{-2 pi i k.R} U_k diag(occ[:, k]) U_k^H, occ (N, Nk)
    weighting the ascending eigenvalues."""
    m = _model(H, executor)
    k = _kpts(ks, m.d)
    w = _weights(ks, kweights, k.shape[1])
    g = np.asfortranarray(np.asarray(occ, dtype=np.float64))
    if g.shape != (m.N, k.shape[1]):
        raise ValueError(f"occupations must have shape (N, Nk) = {(m.N, k.shape[1])}")
    keys = list(rhos.keys())
    nRr = len(keys)
    Rr = np.asfortranarray(np.array(keys, dtype=np.int32).T.reshape(m.d, nRr))
    rho = np.zeros((m.N, m.N, max(nRr, 1)), dtype=np.complex128, order="F")
    _lib.call("lqm_densitymatrix_occ", m.handle, k.shape[1], _vp(k), _vp(w), _vp(g), nRr, _vp(Rr),
              _vp(rho))
    for r, key in enumerate(keys):
        rhos.data[key] = np.array(rho[:, :, r])
    return rhos


def bandloss(H: Hops, ks, target, kweights=None, executor=None):
    """L = sum_k w_k sum_n (e_nk - target_nk)^2 and its cotangent per block t_R,
    rho_g(R) with g = 2 (e - target): two device passes (eigenvalues, then the weighted
    density matrix on the keys of H)."""
    ex = to_executor(executor or "gpu")
    m = _model(H, ex)
    bands, _ = bandmatrix(m, ks)
    Nk = bands.shape[1]
    w = _weights(ks, kweights, Nk)
    wk = np.full(Nk, 1.0 / Nk) if w is None else w
    r = bands - np.asarray(target, dtype=np.float64)
    L = float(np.sum(wk[None, :] * r * r))
    grad = densitymatrix_occ(Hops({R: np.zeros((m.N, m.N), complex) for R in m.keys}), m, ks, 2.0 * r,
                             kweights=kweights)
    return L, grad
