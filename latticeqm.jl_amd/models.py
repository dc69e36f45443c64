"""Input generators for the band path (host-side set-up, run once per model).

These build the exact ``Hops`` the tests and the benchmark feed to the device:

* nearest-neighbour square, triangular and honeycomb models, with the keys and
  blocks that ``Operators.nearestneighbor!`` produces on ``Geometries.square``,
  ``triangular`` and ``honeycomb`` (Operators/nearestneighbor.jl:22-62,
  Operators/gethops.jl:33-140: block (i, j) of key L couples site i of cell L to
  site j of cell 0; the -L block is the adjoint);
* ``addspin`` -- kron(t, sigma_0), spin as the fast index (TightBinding/types.jl:258-270);
* a seeded random dense tight-binding model: complex Gaussian blocks with
  standard deviation exp(-|R|) on a (2 rmax + 1)^d key box, Hermitian by
  construction (t_{-R} = t_R^H, t_0 = (G + G^H)/2).
"""
from __future__ import annotations

import itertools

import numpy as np

from .hops import Hops


def square_nn(t: float = -1.0) -> Hops:
    """1 orbital, keys +-(1,0), +-(0,1): e(k) = 2t(cos 2pi k1 + cos 2pi k2)."""
    one = np.array([[t]], dtype=np.complex128)
    return Hops({(1, 0): one, (-1, 0): one, (0, 1): one, (0, -1): one})


def triangular_nn(t: float = -1.0) -> Hops:
    """1 orbital, A = [(cos30, -sin30), (cos30, sin30)]: neighbours a1, a2, a1 - a2."""
    one = np.array([[t]], dtype=np.complex128)
    keys = [(1, 0), (-1, 0), (0, 1), (0, -1), (1, -1), (-1, 1)]
    return Hops({k: one for k in keys})


def honeycomb_nn(t: float = -1.0) -> Hops:
    """Geometries.honeycomb(): A = sqrt(3)[(cos30, -sin30), (cos30, sin30)], sites
    A at 0 and B at (a1 + a2)/3, bond length 1.  Keys 0, +-(1,0), +-(0,1)."""
    z = np.zeros((2, 2), dtype=np.complex128)
    t0 = z.copy()
    t0[0, 1] = t0[1, 0] = t
    # B in cell (-1,0) / (0,-1) neighbours A in cell 0: block[B, A]
    tm = z.copy()
    tm[1, 0] = t
    h = Hops()
    h[(0, 0)] = t0
    h[(-1, 0)] = tm
    h[(1, 0)] = tm.conj().T
    h[(0, -1)] = tm
    h[(0, 1)] = tm.conj().T
    return h


def addspin(h: Hops, mode: str = "spinhalf") -> Hops:
    """kron(t_R, sigma) per block (TightBinding/types.jl:258-270)."""
    if mode in ("nospin", "id"):
        return h
    if mode in ("spinhalf", "sigma0"):
        s = np.eye(2, dtype=np.complex128)
    elif mode == "sigmax":
        s = np.array([[0, 1], [1, 0]], dtype=np.complex128)
    else:
        raise ValueError(f"unknown spin mode {mode!r}")
    return Hops({R: np.kron(t, s) for R, t in h.items()})


def _cgauss(rng, shape) -> np.ndarray:
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2.0)


def random_tb(N: int, rmax: int = 2, seed: int = 2, dim: int = 2) -> Hops:
    """Random dense tight-binding model on the key box |R_a| <= rmax.

    For R in the half space (first nonzero coordinate positive) t_R is complex
    Gaussian with std exp(-|R|_2); t_{-R} = t_R^H; t_0 = (G + G^H)/2 with G complex
    Gaussian.  Keys are inserted in lexicographic order.
    """
    rng = np.random.default_rng(seed)
    keys = list(itertools.product(range(-rmax, rmax + 1), repeat=dim))
    h = Hops()
    blocks = {}
    for R in keys:
        if all(x == 0 for x in R):
            G = _cgauss(rng, (N, N))
            blocks[R] = (G + G.conj().T) / 2
            continue
        first = next(x for x in R if x != 0)
        if first > 0:
            blocks[R] = _cgauss(rng, (N, N)) * np.exp(-np.linalg.norm(R))
    for R in keys:
        if R in blocks:
            h[R] = blocks[R]
        else:
            mR = tuple(-x for x in R)
            h[R] = blocks[mR].conj().T
    return h


def honeycomb_hubbard(U: float = 3.0, t: float = -1.0):
    """Spinful honeycomb Hubbard model: h = nearest-neighbour t on the three bonds (x) sigma_0
    (N = 4: sites A, B x spin, spin fast), v = U kron(I_2, sigma_x) on the zero key
    (gethubbard(lat; mode=:sigmax, a=0.5, U), Operators/interactions.jl:18-28).  Returns
    (h, v, sublattice labels of the two sites)."""
    h = addspin(honeycomb_nn(t), "spinhalf")
    sx = np.array([[0, 1], [1, 0]], dtype=np.complex128)
    v = Hops({(0, 0): U * np.kron(np.eye(2), sx)})
    return h, v, np.array([0, 1])


def tbg_slater_koster_sparse(m: int = 11, a: float = 1.42, z: float = 3.35, vpppi: float = -2.7,
                             vppsigma: float = 0.48, decay: float = 0.184, cutoff: float = 3.0):
    """tbg_slater_koster as a SparseHops (format = :sparse): the same blocks, built from the
    neighbour pairs within cutoff * a (a k-d tree per key) instead of dense distance arrays, so
    large twist indices (m = 30: 11164 orbitals) fit in host memory."""
    from scipy.spatial import cKDTree
    import scipy.sparse as sp

    from . import lattice
    from .sparse import SparseHops

    lat = lattice.honeycomb_twisted(m, a=a, z=z)
    delta = decay * a
    dmax = cutoff * a
    pos = lat.positions.T  # [n, 3]
    n = pos.shape[0]
    tree = cKDTree(pos)
    out = SparseHops()
    for L in [(0, 0), (1, 0), (0, 1), (1, 1), (1, -1)]:
        shift = lat.A @ np.asarray(L, dtype=float)
        pairs = cKDTree(pos + shift).query_ball_tree(tree, dmax * (1.0 + 1e-9))
        rows = np.repeat(np.arange(n), [len(p) for p in pairs])
        cols = np.fromiter((j for p in pairs for j in p), dtype=np.int64, count=rows.size)
        dr = (pos[rows] + shift - pos[cols]).T
        D = np.sqrt(np.sum(dr * dr, axis=0))
        ok = (D > 1e-6) & (D <= dmax)
        rows, cols, dr, D = rows[ok], cols[ok], dr[:, ok], D[ok]
        c2 = dr[2] ** 2 / D ** 2
        v = vpppi * np.exp(-(D - a) / delta) * (1.0 - c2) + vppsigma * np.exp(-(D - z) / delta) * c2
        keep = v != 0.0
        M = sp.csr_matrix((v[keep].astype(np.complex128), (rows[keep], cols[keep])), shape=(n, n))
        if M.nnz == 0:
            continue
        out[L] = M
        if L != (0, 0):
            out[tuple(-x for x in L)] = M.conj().T
    return out


def tbg_slater_koster(m: int = 11, a: float = 1.42, z: float = 3.35, vpppi: float = -2.7,
                      vppsigma: float = 0.48, decay: float = 0.184, cutoff: float = 3.0) -> Hops:
    """Twisted bilayer graphene of twist index m (4(3m^2 + 3m + 1) orbitals; m = 11: 1588)
    from honeycomb_twisted(m; a, z) (Structure/Geometries/2d.jl:222-227,
    Structure/Lattices/twist.jl:45-173) with the two-centre Slater-Koster hopping
    t(d) = vpppi e^{-(|d| - a)/delta} (1 - c^2) + vppsigma e^{-(|d| - z)/delta} c^2,
    c = d_z/|d|, delta = decay * a, zero beyond cutoff * a and on site, over the images
    R in {-1, 0, 1}^2 (empty blocks trimmed).  a is the C-C distance, as the lattice
    generator takes it."""
    from . import lattice

    lat = lattice.honeycomb_twisted(m, a=a, z=z)
    delta = decay * a
    dmax = cutoff * a

    def t(dr):
        D = np.sqrt(np.sum(dr * dr, axis=0))
        ok = (D > 1e-6) & (D <= dmax)
        Ds = np.where(ok, D, 1.0)
        c2 = dr[2] ** 2 / Ds ** 2
        v = vpppi * np.exp(-(Ds - a) / delta) * (1.0 - c2) + vppsigma * np.exp(-(Ds - z) / delta) * c2
        return np.where(ok, v, 0.0)

    t.vectorized = True
    half = [(0, 0), (1, 0), (0, 1), (1, 1), (1, -1)]
    return lattice.gethops(lat, t, neighbors=half, precision=0.0)
