"""The allocation-free DOS entry point against the workspace one, through the C ABI.

lqm_dos_from_bands_dev_noalloc runs the same kernels on the same slices as
lqm_dos_from_bands_dev, with the partials in the library's fixed device array instead of the
context workspace, so the two agree bit for bit.  Two calls in a row, the second on a smaller
input (the array keeps the first call's partials), then the size limit.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dos_pair(torch, L, ctx, bands, kw, om, gamma):
    """(workspace result, allocation-free result) for bands (Nk, N) on the device."""
    from latticeqm_amd import _lib

    vp = ctypes.c_void_p
    Nk, N = bands.shape
    Nw = om.shape[0]
    out_ws = torch.full((Nw,), float("nan"), dtype=torch.float64, device="cuda")
    out_na = torch.full((Nw,), float("nan"), dtype=torch.float64, device="cuda")
    kwp = vp(kw.data_ptr()) if kw is not None else None
    _lib.check(L.lqm_dos_from_bands_dev(ctx, N, Nk, vp(bands.data_ptr()), kwp, Nw, vp(om.data_ptr()), gamma,
                                        vp(out_ws.data_ptr())), "lqm_dos_from_bands_dev")
    _lib.check(L.lqm_dos_from_bands_dev_noalloc(ctx, N, Nk, vp(bands.data_ptr()), kwp, Nw, vp(om.data_ptr()),
                                                gamma, vp(out_na.data_ptr())), "lqm_dos_from_bands_dev_noalloc")
    return out_ws.cpu().numpy(), out_na.cpu().numpy()


def test_noalloc_dos_matches_workspace_entry_point(gpu):
    import torch
    from latticeqm_amd import _lib

    L = _lib.lib()
    rng = np.random.default_rng(7)
    # first call: 64 bands x 4096 k-points, 1500 frequencies (6 frequency tiles, 170 slices)
    b1 = torch.from_numpy(np.sort(rng.normal(size=(4096, 64)), axis=1)).cuda()
    w1 = torch.from_numpy(rng.uniform(0.5, 1.5, 4096) / 4096).cuda()
    om1 = torch.linspace(-4.0, 4.0, 1500, dtype=torch.float64, device="cuda")
    ws1, na1 = _dos_pair(torch, L, gpu.handle, b1, w1, om1, 0.02)
    assert np.all(np.isfinite(na1))
    assert np.array_equal(ws1, na1)
    # second call, smaller: 8 bands x 300 k-points, 333 frequencies, uniform weights
    b2 = torch.from_numpy(np.sort(rng.normal(size=(300, 8)), axis=1)).cuda()
    om2 = torch.linspace(-3.0, 3.0, 333, dtype=torch.float64, device="cuda")
    ws2, na2 = _dos_pair(torch, L, gpu.handle, b2, None, om2, 0.05)
    assert np.all(np.isfinite(na2))
    assert np.array_equal(ws2, na2)
    # and a closed form: one band at 0 on one k-point is gamma / (w^2 + gamma^2)
    b3 = torch.zeros((1, 1), dtype=torch.float64, device="cuda")
    ws3, na3 = _dos_pair(torch, L, gpu.handle, b3, None, om2, 0.05)
    w = om2.cpu().numpy()
    assert np.allclose(na3, 0.05 / (w * w + 0.05 ** 2), rtol=1e-14, atol=0)
    assert np.array_equal(ws3, na3)


def test_noalloc_dos_rejects_too_many_frequencies(gpu):
    from latticeqm_amd import _lib

    L = _lib.lib()
    st = L.lqm_dos_from_bands_dev_noalloc(gpu.handle, 4, 4, ctypes.c_void_p(16), None, 262145,
                                          ctypes.c_void_p(16), 0.1, ctypes.c_void_p(16))
    assert st == _lib.LQM_ERR_UNSUPPORTED
    assert b"262144" in L.lqm_last_error()
