"""Scattered simulation nodes (create_triangulation, SimulationInterpolater.cpp:
105-117; interpolated_value, :33-102) on the CPU: the library's Delaunay
triangulation against scipy's Qhull on the same points, the oracle's lookup
against scipy's LinearNDInterpolator (Qhull's triangulation, barycentric
weights), and the lookup's independence of the bucket layout."""
import numpy as np
import pytest
from scipy.interpolate import LinearNDInterpolator
from scipy.spatial import Delaunay

from jetshow_amd import abi, host, textio
from oracle import oracle

PC = abi.PC


def _tri_set(T):
    return set(tuple(sorted(t)) for t in np.asarray(T).tolist())


def _ccw(z, r, T):
    return (z[T[:, 1]] - z[T[:, 0]]) * (r[T[:, 2]] - r[T[:, 0]]) - (r[T[:, 1]] - r[T[:, 0]]) * (z[T[:, 2]] - z[T[:, 0]])


@pytest.mark.parametrize("n", [7, 60, 800, 20000])
def test_triangulation_equals_qhull(n):
    rng = np.random.default_rng(n)
    z = rng.uniform(0.0, 20.0, n)
    r = rng.uniform(0.0, 0.5, n)
    T = textio.triangulate(z, r)
    S = Delaunay(np.stack([z, r], 1)).simplices
    assert _tri_set(T) == _tri_set(S)
    assert (_ccw(z, r, T) > 0).all()
    # canonical form: each triangle starts at its smallest index, rows sorted
    assert (T[:, 0] < T[:, 1]).all() and (T[:, 0] < T[:, 2]).all()
    assert [tuple(t) for t in T.tolist()] == sorted(tuple(t) for t in T.tolist())


def test_triangulation_jittered_grid_with_straight_hull():
    """nodes of a jittered (z, r) grid whose axis nodes lie exactly on r = 0:
    the hull's straight stretches are covered completely (no slivers)"""
    rng = np.random.default_rng(2)
    zz, rr = np.meshgrid(np.linspace(0, 20, 50), np.linspace(0, 0.5, 30), indexing="ij")
    z = zz.ravel() + rng.normal(0, 1e-3, zz.size)
    r = np.abs(rr.ravel() + rng.normal(0, 1e-4, zz.size))
    r[::30] = 0.0
    T = textio.triangulate(z, r)
    assert _tri_set(T) == _tri_set(Delaunay(np.stack([z, r], 1)).simplices)


def test_triangulation_of_cocircular_regular_grid_is_valid():
    """a regular grid (every cell's corners on one circle, so the Delaunay
    triangulation is not unique): a valid triangulation, two triangles per
    cell, covering the box exactly"""
    nz, nr = 23, 11
    zz, rr = np.meshgrid(np.linspace(0, 20, nz), np.linspace(0, 0.5, nr), indexing="ij")
    z, r = zz.ravel(), rr.ravel()
    T = textio.triangulate(z, r)
    a = _ccw(z, r, T)
    assert len(T) == 2 * (nz - 1) * (nr - 1)
    assert (a > 0).all()
    assert abs(0.5 * a.sum() - 20 * 0.5) < 1e-12


def test_triangulation_duplicates_and_degenerate_sets():
    rng = np.random.default_rng(4)
    z = rng.uniform(0, 1, 40)
    r = rng.uniform(0, 1, 40)
    zd = np.concatenate([z, z[:10]])
    rd = np.concatenate([r, r[:10]])
    assert _tri_set(textio.triangulate(zd, rd)) == _tri_set(textio.triangulate(z, r))
    with pytest.raises(host.JsxError):
        textio.triangulate(np.linspace(0, 1, 10), 2 * np.linspace(0, 1, 10))   # one line
    with pytest.raises(host.JsxError):
        textio.triangulate([0.0, 1.0], [0.0, 1.0])


def _scattered(n, seed, nan_value=0.0):
    rng = np.random.default_rng(seed)
    z = np.concatenate([rng.uniform(0, 20, n), [0, 20, 0, 20]])
    r = np.concatenate([rng.uniform(0, 0.5, n), [0, 0, 0.5, 0.5]])
    v = np.sin(0.7 * z) * np.exp(-3 * r) + 0.05 * rng.normal(size=z.size)
    return z, r, v, oracle.TriGrid(z, r, v, textio.triangulate(z, r), nan_value)


def _queries(m, seed):
    rng = np.random.default_rng(seed)
    qz = rng.uniform(-1, 21, m)
    qr = rng.uniform(0, 0.6, m)
    ph = rng.uniform(0, 2 * np.pi, m)
    return np.stack([qr * np.cos(ph) * PC, qr * np.sin(ph) * PC, qz * PC], 1), qz, qr


def test_oracle_lookup_matches_scipy_linear_interpolation():
    z, r, v, g = _scattered(3000, 5, nan_value=-7.0)
    pts, qz, qr = _queries(6000, 6)
    a = oracle.tri_grid_interp(g, pts)
    ref = LinearNDInterpolator(np.stack([z, r], 1), v, fill_value=-7.0)(
        np.stack([qz, np.hypot(pts[:, 0], pts[:, 1]) / PC], 1))
    inside = ref != -7.0
    assert inside.sum() > 4000 and (~inside).sum() > 500
    assert np.array_equal(a[~inside], ref[~inside])
    assert np.abs(a[inside] - ref[inside]).max() <= 1e-13 * np.abs(v).max()
    # at the nodes: the node values
    at = oracle.tri_grid_interp(g, np.stack([r * PC, np.zeros_like(r), z * PC], 1))
    assert np.abs(at - v).max() <= 1e-13 * np.abs(v).max()


def test_oracle_lookup_independent_of_bucket_layout():
    """the lookup takes the lowest-numbered triangle holding the point, so any
    bucket layout gives the same bits -- including for queries on edges"""
    z, r, v, g = _scattered(500, 7)
    T = g.tri
    # edge midpoints and nodes: points two or more triangles hold
    mids = 0.5 * (np.stack([z[T[:, 0]], r[T[:, 0]]], 1) + np.stack([z[T[:, 1]], r[T[:, 1]]], 1))
    pts = np.concatenate([np.stack([mids[:, 1] * PC, 0 * mids[:, 0], mids[:, 0] * PC], 1),
                          np.stack([r * PC, 0 * r, z * PC], 1), _queries(2000, 8)[0]])
    a = oracle.tri_grid_interp(g, pts)
    g2 = oracle.TriGrid(z, r, v, T, 0.0)
    g2.nbz, g2.nbr = 1, 1
    g2.bdz, g2.bdr = g2.bzl - g2.bz0, g2.brl - g2.br0
    g2.boff = np.array([0, len(T)], dtype=np.int32)
    g2.blist = np.arange(len(T), dtype=np.int32)
    b = oracle.tri_grid_interp(g2, pts)
    assert np.array_equal(a, b)
