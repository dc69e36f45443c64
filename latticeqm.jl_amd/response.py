"""Kubo linear response on the GPU band path (LatticeQM.jl's LinearResponse module).

=====================================  ========================================================
this module                            reference
=====================================  ========================================================
``opticalconductivity(w, H, J1, J2,    opticalconductivity(frequencies, H, J1, J2, ks; mu, G, T)
ks)``                                  (LinearResponse/opticalconductivity.jl:101-127) with
                                       kubo! (kubo.jl:11-31); ``klin`` instead of ``ks``: the
                                       regular klin x klin grid (:65-68)
``opticalconductivity(w, i, j, H,      opticalconductivity(frequencies, i, j, H, lat)
lat)``                                 (:50-53): currents from getcurrentoperators(lat, H)
``opticalconductivityXX/XY``           (:14-30)
``sigma_from_chi``                     sigma(w) = -(chi(w) - chi(w_0)) / (w + iG), the recipe of
                                       test/test_linearresponse.jl:10
=====================================  ========================================================

All of it returns the Kubo current-current correlator

    chi(w) = -(i/Nk) sum_k sum_{m != n} (f_m - f_n) <m|J1|n><n|J2^H|m> / (e_m - e_n - w - iG),

f = fermidirac(e; mu, T), computed by ``lqm_kubo``: the eigenvectors of the whole grid from the
batched eigensolver, U^H J U on the matrix cores, the pair sums over all frequencies in one
kernel.  Cartesian directions i, j are 0-based here (x = 0, y = 1).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .hops import Hops, regulargrid


def _pop(kw, name, alias, default):
    if name in kw:
        return kw.pop(name)
    return kw.pop(alias, default)


def opticalconductivity(frequencies, *args, klin: int | None = None, kweights=None, executor=None, **kw):
    """chi(w) on the frequencies; keywords mu (μ) = 0, Gamma (Γ) = 0.025, T = 0.1."""
    from .api import GPUModel, _kpts, _vp, _weights, to_executor
    from .lattice import Lattice, getcurrentoperators

    mu = float(_pop(kw, "mu", "μ", 0.0))
    gamma = float(_pop(kw, "Gamma", "Γ", 0.025))
    T = float(kw.pop("T", 0.1))
    if kw:
        raise TypeError(f"unexpected keywords {sorted(kw)}")
    if len(args) >= 4 and isinstance(args[0], (int, np.integer)) and isinstance(args[3], Lattice):
        i, j, H, lat = args[:4]
        J = getcurrentoperators(lat, H)
        rest = args[4:]
        J1, J2 = J[i], J[j]
    elif len(args) >= 3:
        H, J1, J2 = args[:3]
        rest = args[3:]
    else:
        raise TypeError("opticalconductivity(w, H, J1, J2[, ks]) or opticalconductivity(w, i, j, H, lat[, ks])")
    if rest:
        ks = rest[0]
    elif klin is not None:
        ks = regulargrid(nk=klin ** 2)
    else:
        raise ValueError("give the k-points or klin")
    om = np.atleast_1d(np.ascontiguousarray(frequencies, dtype=np.float64))
    if om.size == 0:
        raise ValueError("at least one frequency")
    ex = H.exec if isinstance(H, GPUModel) else to_executor(executor or "gpu")
    mh = H if isinstance(H, GPUModel) else GPUModel(H, ex)
    m1 = J1 if isinstance(J1, GPUModel) else GPUModel(J1, ex)
    m2 = m1 if J2 is J1 else (J2 if isinstance(J2, GPUModel) else GPUModel(J2, ex))
    k = _kpts(ks, mh.d)
    w = _weights(ks, kweights, k.shape[1])
    out = np.zeros(2 * om.size)
    _lib.call("lqm_kubo", mh.handle, m1.handle, m2.handle, k.shape[1], _vp(k), _vp(w), om.size, _vp(om), mu, gamma,
              T, _vp(out))
    return out[0::2] + 1j * out[1::2]


def opticalconductivityXX(frequencies, H: Hops, lat, *args, **kw):
    """chi_xx (the raw correlator, not sigma)."""
    return opticalconductivity(frequencies, 0, 0, H, lat, *args, **kw)


def opticalconductivityXY(frequencies, H: Hops, lat, *args, **kw):
    """chi_xy (the raw correlator, not sigma)."""
    return opticalconductivity(frequencies, 0, 1, H, lat, *args, **kw)


def sigma_from_chi(chi, frequencies, Gamma: float) -> np.ndarray:
    """sigma(w) = -(chi(w) - chi(w_0)) / (w + iG): the static piece (the correlator at the first
    frequency) subtracted and the photon energy divided out."""
    chi = np.asarray(chi)
    return -(chi - chi[0]) / (np.asarray(frequencies) + 1j * Gamma)
