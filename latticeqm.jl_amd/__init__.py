"""MI355X (gfx950) k-space band path of LatticeQM.jl.

Bloch assembly, batched Hermitian eigensolve, DOS, chemical potential and
density-matrix accumulation over whole k-grids, computed by hand-written HIP
kernels in liblqmgpu.so (C ABI: include/lqm_gpu.h).

The directory name contains a dot, so it is loaded by path; the loaders in this
repository register it as ``latticeqm_amd``.
"""
from .hops import Hops, Mesh, kpath, points, regulargrid  # noqa: F401
from . import models  # noqa: F401
from ._lib import header_symbols  # noqa: F401


def __getattr__(name):
    # The compute API needs liblqmgpu.so; import it lazily so that the host-side
    # containers stay importable where the library has not been built.
    import importlib

    if name in ("api", "meanfield", "bdg", "berry", "response", "sparse"):
        return importlib.import_module(__name__ + "." + name)
    api = importlib.import_module(__name__ + ".api")
    if hasattr(api, name):
        return getattr(api, name)
    raise AttributeError(name)
