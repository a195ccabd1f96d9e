"""The CPU oracle's multi-threaded schedule (the CPU baseline) and its IC(0):
the OpenMP factorization equals the sequential oracle bit for bit for every
thread count, the OpenMP PCG follows the sequential one, and IC(0) reproduces
the matrix on its own pattern (its defining property) and is exact where no
fill can occur."""
import numpy as np
import pytest


def _systems(wl):
    out = {
        "grid3d_log": wl.grid_laplacian((14, 12, 10), "loguniform", seed=7),
        "grid2d_unit": wl.grid_laplacian((40, 30)),
        "rgg": wl.random_geometric_graph(3000, seed=5),
        "two_components": None,
    }
    n, r, c, v = wl.grid_laplacian((9, 9), "uniform_0.5_2", seed=1)
    out["two_components"] = (2 * n, np.concatenate([r, r + n]), np.concatenate([c, c + n]), np.concatenate([v, v]))
    return out


@pytest.mark.parametrize("name", ["grid3d_log", "grid2d_unit", "rgg", "two_components", "sddm", "star"])
def test_factor_omp_equals_sequential_oracle(orc, wl, name):
    if name == "sddm":
        n, r, c, v = wl.grid_sddm((12, 11, 10))
        fac, _, _, _ = orc.prepare_system(n, r, c, v, kind="sddm", order="nnz-sort", order_seed=3)
    elif name == "star":
        m = 400                                  # one column gathers every leaf
        u = np.zeros(m, np.int64)
        w = np.arange(1, m + 1)
        n, r, c, v = wl.laplacian_triplets(m + 1, u, w, np.linspace(0.5, 2.0, m))
        fac, _, _, _ = orc.prepare_system(n, r, c, v, order="random", order_seed=2)
    else:
        n, r, c, v = _systems(wl)[name]
        fac, _, _, _ = orc.prepare_system(n, r, c, v, order="nnz-sort", order_seed=1)
    f = orc.factor(fac, seed=9)
    for threads in (1, 3, 8):
        g = orc.factor_omp(fac, seed=9, threads=threads)
        np.testing.assert_array_equal(g.ptr, f.ptr)
        np.testing.assert_array_equal(g.ind, f.ind)
        np.testing.assert_array_equal(g.val, f.val)
        np.testing.assert_array_equal(g.diag, f.diag)
        np.testing.assert_array_equal(g.level, f.level)
        assert g.stats == f.stats


def test_factor_omp_catches_a_wrong_seed(orc, wl):
    n, r, c, v = wl.grid_laplacian((10, 10), "loguniform", seed=3)
    fac, _, _, _ = orc.prepare_system(n, r, c, v)
    assert not np.array_equal(orc.factor_omp(fac, seed=1, threads=4).val, orc.factor(fac, seed=2).val)


@pytest.mark.parametrize("name", ["grid3d_log", "grid2d_unit", "sddm"])
def test_pcg_omp_follows_sequential_pcg(orc, wl, name):
    if name == "sddm":
        n, r, c, v = wl.grid_sddm((14, 14, 14))
        fac, sys_, perm, app = orc.prepare_system(n, r, c, v, kind="sddm", order="nnz-sort", order_seed=1)
    else:
        n, r, c, v = _systems(wl)[name]
        fac, sys_, perm, app = orc.prepare_system(n, r, c, v, order="nnz-sort", order_seed=1)
    f = orc.factor(fac, seed=1)
    if app:
        f = orc.factor_trim_last(f)
    b = orc.normal_vector(42, n)
    if name != "sddm":
        b -= b.mean()
    x, st = orc.pcg(sys_, f, b[perm], tol=1e-8)
    for threads in (1, 4):
        y, so = orc.pcg_omp(sys_, f, b[perm], tol=1e-8, threads=threads)
        assert so["status"] == 0 and so["converged"] == 1
        assert abs(so["iterations"] - st["iterations"]) <= 1, (so, st)
        assert np.linalg.norm(y - x) <= 1e-7 * np.linalg.norm(x)
        assert so["rel_res_true"] < 1e-6


def test_normal_vector_matches_scalar_draws(orc):
    v = orc.normal_vector(42, 1000, 3)
    np.testing.assert_array_equal(v, [orc.normal(42, i, 3) for i in range(1000)])


# ------------------------------------------------------------------ IC(0)
def _ldl(f):
    L = f.dense_L()
    return L @ np.diag(f.diag) @ L.T


@pytest.mark.parametrize("dims", [(12, 10), (6, 5, 4), "rgg"])
def test_ic0_reproduces_the_matrix_on_its_pattern(orc, wl, dims):
    if dims == "rgg":
        n, r, c, v = wl.random_geometric_graph(150, seed=3)
    else:
        n, r, c, v = wl.grid_laplacian(dims, "loguniform", seed=4)
    fac, _, _, _ = orc.prepare_system(n, r, c, v, order="nnz-sort", order_seed=1)
    f = orc.ic0(fac)
    A = fac.dense()
    M = _ldl(f)
    pat = A != 0
    # (L D L^T)_ij = a_ij on the pattern (a pivot clamped to 0 would be the
    # one exception; the dropped fill keeps every pivot of these grids positive)
    np.testing.assert_allclose(M[pat], A[pat], rtol=1e-10, atol=1e-12 * np.abs(A).max())
    assert np.any(np.abs(M[~pat]) > 1e-12)       # incomplete: fill outside the pattern is dropped
    assert np.all(f.diag[np.diag(A) > 0] > 0)      # isolated vertices keep a zero pivot
    # the forward levels increase along every entry of L
    for k in range(n):
        for p in range(f.ptr[k], f.ptr[k + 1]):
            assert f.level[f.ind[p]] > f.level[k]


def test_ic0_is_exact_on_a_tree(orc, wl):
    # a path (and any tree eliminated leaf-first) has no fill: IC(0) = LDL^T
    m = 50
    u = np.arange(m - 1)
    n, r, c, v = wl.laplacian_triplets(m, u, u + 1, np.linspace(1.0, 3.0, m - 1))
    fac, _, _, _ = orc.prepare_system(n, r, c, v, order="as-is")
    f = orc.ic0(fac)
    np.testing.assert_allclose(_ldl(f), fac.dense(), atol=1e-12)


def test_ic0_pin_catches_a_dropped_pivot(orc, wl):
    """The pattern pin fails on an oracle whose update sums leave out d_m
    (a graph with triangles: on a grid two neighbours share no lower
    neighbour, so those sums are empty)."""
    n, r, c, v = wl.random_geometric_graph(150, seed=3)
    fac, _, _, _ = orc.prepare_system(n, r, c, v, order="nnz-sort", order_seed=1)
    A = fac.dense()
    pat = A != 0
    good = _ldl(orc.ic0(fac))
    bad = _ldl(orc.ic0(fac, mutation=orc.MUT_IC0_NO_PIVOT))
    assert np.allclose(good[pat], A[pat], rtol=1e-10, atol=1e-9)
    assert not np.allclose(bad[pat], A[pat], rtol=1e-6, atol=1e-9)
