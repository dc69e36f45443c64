"""The C-ABI library loads and exports every function include/lqm_gpu.h declares.
No compute calls (CPU only)."""
import ctypes
import os


def test_header_declares_core_entry_points(pkg):
    syms = pkg.header_symbols()
    for name in ["lqm_context_create", "lqm_model_create", "lqm_bandmatrix", "lqm_densitymatrix",
                 "lqm_dos", "lqm_chemicalpotential", "lqm_last_error"]:
        assert name in syms


def test_library_exports_every_declared_symbol(pkg):
    from latticeqm_amd import _lib

    assert os.path.exists(_lib.LIB_PATH), "build liblqmgpu.so first (__graft_entry__.build())"
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in pkg.header_symbols() if not hasattr(L, s)]
    assert not missing, missing
    # every exported prototype of the binding is declared in the header
    assert set(_lib._PROTOS) <= set(pkg.header_symbols())


def test_abi_version_and_error_string():
    from latticeqm_amd import _lib

    L = _lib.lib()
    assert L.lqm_abi_version() >= 1
    assert isinstance(L.lqm_last_error(), bytes)
