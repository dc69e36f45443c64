"""GPU bands of twisted bilayer graphene against the reference's TBG fixtures and the
FP64 oracle (eigenvalue bar 1e-10 relative to max(|e|, ||H(k)||))."""
import importlib

import numpy as np
import pytest

from tbg_common import FIXTURES, check

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lat(pkg):
    return importlib.import_module(pkg.__name__ + ".lattice")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n", [1, 2, 3])
def test_tbg_fixtures_on_gpu(pkg, oracle, gpu, lat, n, precision):
    fx = FIXTURES[f"n{n}"]
    h = lat.graphene(lat.honeycomb_twisted(n), tz=0.32)
    ex = pkg.api.to_executor("gpu", precision=precision)
    ks = np.array([[0.0, 1 / 3], [0.0, -1 / 3]])
    b, _ = pkg.api.bandmatrix(h, ks, executor=ex)
    check(b[:, 0], fx["gamma"], fx["tol"])
    if "K" in fx:
        check(b[:, 1], fx["K"], fx["tol"])
    # and a k-grid against the oracle
    grid = pkg.regulargrid(nk=16)
    R, T = h.pack()
    ref = oracle.bands(R, T, grid)
    bg, _ = pkg.api.bandmatrix(h, grid, executor=ex)
    scale = np.array([np.linalg.norm(h(k), 2) for k in grid.T])
    assert np.max(np.abs(bg - ref) / np.maximum(np.abs(ref), scale[None, :])) < 1e-10
