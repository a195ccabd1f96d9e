"""Synthetic workloads: the graph Laplacians and SDDM systems the benchmark and
the tests run on (the reference ships no data; its Julia preprocessing
generates grids with Laplacians.jl, write_graph.jl:256-289).

All generators return Matrix-Market-style triplets (0-based, both triangles)
that go through the same input path as a .mtx file.
"""
from __future__ import annotations

import numpy as np


def grid_edges(dims):
    """Undirected nearest-neighbour edges (i < j) of a 2D/3D grid, C order."""
    dims = tuple(int(d) for d in dims)
    n = int(np.prod(dims))
    idx = np.arange(n, dtype=np.int64).reshape(dims)
    us, vs = [], []
    for ax in range(len(dims)):
        a = np.take(idx, range(dims[ax] - 1), axis=ax).ravel()
        b = np.take(idx, range(1, dims[ax]), axis=ax).ravel()
        us.append(a)
        vs.append(b)
    return n, np.concatenate(us), np.concatenate(vs)


def laplacian_triplets(n, u, v, w):
    """Off-diagonal adjacency triplets -w in both triangles; the input path
    appends the diagonal as the column sum (pre_process.hpp:210-231)."""
    rows = np.concatenate([u, v]).astype(np.int32)
    cols = np.concatenate([v, u]).astype(np.int32)
    vals = -np.concatenate([w, w]).astype(np.float64)
    return n, rows, cols, vals


def grid_laplacian(dims, weights="unit", seed=7):
    """Weighted grid Laplacian.  weights: 'unit', 'loguniform' (10^U(-3,3)),
    or 'uniform_0.5_2' (U[0.5, 2]).  Weights drawn with numpy PCG64(seed)."""
    n, u, v = grid_edges(dims)
    m = len(u)
    if weights == "unit":
        w = np.ones(m)
    elif weights == "loguniform":
        w = 10.0 ** np.random.default_rng(seed).uniform(-3.0, 3.0, m)
    elif weights == "uniform_0.5_2":
        w = np.random.default_rng(seed).uniform(0.5, 2.0, m)
    else:
        raise ValueError(weights)
    return laplacian_triplets(n, u, v, w)


def grid_sddm(dims, shift=1e-2):
    """7-point (3D) / 5-point (2D) Poisson SDDM with diagonal = degree + shift."""
    n, u, v = grid_edges(dims)
    deg = np.bincount(u, minlength=n) + np.bincount(v, minlength=n)
    rows = np.concatenate([u, v, np.arange(n)]).astype(np.int32)
    cols = np.concatenate([v, u, np.arange(n)]).astype(np.int32)
    vals = np.concatenate([-np.ones(2 * len(u)), deg.astype(np.float64) + shift])
    return n, rows, cols, vals


def random_geometric_graph(n, seed, mean_degree=8.0):
    """RGG in the unit square, radius sqrt(mean_degree / (pi n)), weight 1/dist."""
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    pts = rng.random((n, 2))
    r = np.sqrt(mean_degree / (np.pi * n))
    pairs = cKDTree(pts).query_pairs(r, output_type="ndarray")
    u, v = pairs[:, 0], pairs[:, 1]
    d = np.linalg.norm(pts[u] - pts[v], axis=1)
    d = np.maximum(d, 1e-12)
    return laplacian_triplets(n, u, v, 1.0 / d)


def dense_from_triplets(n, rows, cols, vals):
    """Dense Laplacian as the input path builds it (tests only, small n)."""
    a = np.zeros((n, n))
    has_diag = np.any(rows == cols)
    off = rows != cols
    np.add.at(a, (rows[off], cols[off]), -np.abs(vals[off]))
    if has_diag:
        d = rows == cols
        np.add.at(a, (rows[d], cols[d]), np.abs(vals[d]))
    else:
        a[np.arange(n), np.arange(n)] = -a.sum(axis=0)
    return a
