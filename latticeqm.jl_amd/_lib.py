"""ctypes binding of liblqmgpu.so (the C ABI declared in include/lqm_gpu.h).

The library is loaded from this package directory.  There is no fallback: if it
has not been built the import of any compute function raises, naming the build
step.  Status codes map to the reference's error conventions:
LQM_ERR_ARG -> ValueError (DimensionMismatch/ArgumentError), LQM_ERR_NOCONV ->
RuntimeError (LAPACKException analogue), LQM_ERR_NONFINITE -> FloatingPointError,
others -> RuntimeError.
"""
from __future__ import annotations

import ctypes
import os
import re

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "liblqmgpu.so")
HEADER = os.path.join(os.path.dirname(PKG_DIR), "include", "lqm_gpu.h")

LQM_OK, LQM_ERR_ARG, LQM_ERR_DEVICE, LQM_ERR_UNSUPPORTED, LQM_ERR_NOCONV, LQM_ERR_NONFINITE = range(6)
LQM_PREC_FP64 = 0
LQM_PREC_FP32_REFINED = 1

_lib = None

P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
PP = ctypes.POINTER(ctypes.c_void_p)

# name -> argtypes (every function returns int status unless listed in _RESTYPE)
_PROTOS = {
    "lqm_last_error": [],
    "lqm_abi_version": [],
    "lqm_context_create": [I, I, PP],
    "lqm_context_destroy": [P],
    "lqm_context_set_stream": [P, P],
    "lqm_context_synchronize": [P],
    "lqm_context_set_timing": [P, I],
    "lqm_context_set_graphs": [P, I],
    "lqm_context_set_progress": [P, P, P],
    "lqm_context_save_tuning": [P, ctypes.c_char_p],
    "lqm_context_load_tuning": [P, ctypes.c_char_p],
    "lqm_context_stage_times": [P, P, P, P, I],
    "lqm_stage_name": [I],
    "lqm_context_set_variant": [P, I, I],
    "lqm_context_set_refinement": [P, I],
    "lqm_context_refinement_report": [P, P],
    "lqm_context_last_precision": [P, P],
    "lqm_context_set_refinement_threshold": [P, D],
    "lqm_context_get_variant": [P, I, I, P],
    "lqm_autotune": [P, I, P],
    "lqm_model_create": [P, I, I, I, P, P, PP],
    "lqm_model_create_dev": [P, I, I, I, P, P, PP],
    "lqm_model_destroy": [P],
    "lqm_fouriersum": [P, I, P, P],
    "lqm_fouriersum_dev": [P, I, P, P],
    "lqm_bandmatrix": [P, I, P, P, P],
    "lqm_bandmatrix_dev": [P, I, P, P, P],
    "lqm_dos": [P, I, P, P, I, P, D, P],
    "lqm_ldos": [P, I, P, P, I, P, D, P],
    "lqm_ldos_dev": [P, I, P, P, I, P, D, P],
    "lqm_density": [P, I, P, P, D, P],
    "lqm_density_dev": [P, I, P, P, D, P],
    "lqm_kubo": [P, P, P, I, P, P, I, P, D, D, D, P],
    "lqm_kubo_dev": [P, P, P, I, P, P, I, P, D, D, D, P],
    "lqm_sparse_model_create": [P, I, I, I, P, ctypes.c_int64, P, P, P, P, PP],
    "lqm_sparse_model_destroy": [P],
    "lqm_sparse_model_bound": [P, P],
    "lqm_eigen_partial": [P, I, P, I, I, D, D, I, P, P, P],
    "lqm_eigen_partial_dev": [P, I, P, I, I, D, D, I, P, P, P],
    "lqm_dos_from_bands_dev": [P, I, I, P, P, I, P, D, P],
    "lqm_dos_from_bands_dev_noalloc": [P, I, I, P, P, I, P, D, P],
    "lqm_chemicalpotential": [P, I, P, P, D, D, P],
    "lqm_chemicalpotential_from_bands_dev": [P, I, I, P, P, D, D, P],
    "lqm_densitymatrix": [P, I, P, P, D, D, I, P, P, P],
    "lqm_densitymatrix_dev": [P, I, P, P, D, D, I, P, P, P, P],
    "lqm_bandmatrix_expval": [P, I, P, I, P, P, P],
    "lqm_bandmatrix_expval_dev": [P, I, P, I, P, P, P],
    "lqm_grandpotential": [P, I, P, P, D, D, P, P],
    "lqm_grandpotential_dev": [P, I, P, P, D, D, P, P],
    "lqm_densitymatrix_occ": [P, I, P, P, P, I, P, P],
    "lqm_densitymatrix_occ_dev": [P, I, P, P, P, I, P, P, P],
    "lqm_hartreefock_scf_batch": [P, I, I, P, D, D, D, I, I, I],
    "lqm_bdg_densitymatrix": [P, I, P, P, D, D, I, P, P, P],
    "lqm_link_overlaps": [P, I, I, I, P, I, P, P],
    "lqm_link_phases": [P, I, I, I, P, I, P, P],
    "lqm_link_overlaps_dev": [P, I, I, P, I, P, P],
    "lqm_link_phases_dev": [P, I, I, P, I, P, P],
    "lqm_hartreefock_grandpotential": [P, I, I, I, P, P, I, P, P, I, P, P, I, I, P, P, D, D, P, P, P],
    "lqm_bdg_filling": [P, I, P, P, D, D, P],
    "lqm_bdg_chemicalpotential": [P, I, P, P, D, D, D, D, P, P],
    "lqm_hartreefock_meanfield": [P, I, I, I, P, P, I, P, P, I, P, P, I, P, P, P, P],
    "lqm_hartreefock_scf": [P, I, I, I, P, P, I, P, P, I, P, P, I, P, P, D, D, D, D, I, I, P, P, P,
                            P, P, P],
}
_RESTYPE = {"lqm_last_error": ctypes.c_char_p, "lqm_abi_version": I, "lqm_stage_name": ctypes.c_char_p}
NSTAGES = 28


class HFProblem(ctypes.Structure):
    """lqm_hf_problem of include/lqm_gpu.h."""
    _fields_ = [("N", I), ("nRh", I), ("Rh", P), ("Th", P), ("nRv", I), ("Rv", P), ("Tv", P),
                ("nRr", I), ("Rr", P), ("rho", P), ("Nk", I), ("ks", P), ("kweights", P),
                ("filling", D), ("result", D * 8), ("converged", I), ("status", I)]
LQM_HF_HARTREE = 1
LQM_HF_FOCK = 2


def header_symbols(path: str = HEADER) -> list[str]:
    """Names of the functions declared in include/lqm_gpu.h."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lqm_[a-z0-9_]+)\s*\(", text)))


def lib():
    """Load liblqmgpu.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    for name, args in _PROTOS.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = _RESTYPE.get(name, I)
    _lib = L
    return L


class LQMError(RuntimeError):
    pass


def check(status: int, what: str = "") -> None:
    if status == LQM_OK:
        return
    msg = (lib().lqm_last_error() or b"").decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if status == LQM_ERR_ARG:
        raise ValueError(text)
    if status == LQM_ERR_NONFINITE:
        raise FloatingPointError(text)
    raise LQMError(f"[status {status}] {text}")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)

# fn(problem, iterations, converged, user) of lqm_context_set_progress
PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p)
