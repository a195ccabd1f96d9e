"""Pins of the CPU oracle (oracle/rchol_oracle.c) against known answers,
closed forms, library routines and invariants.  CPU only.

Every function of the oracle is checked by something a plausible mistake would
fail (a dropped term, a wrong index, a flipped sign); the `mutation` tests
plant such mistakes in the oracle and assert that the pins catch them.
"""
import numpy as np
import pytest

from conftest import golden_path, read_mtx_py


# ------------------------------------------------------------------ RNG
def test_philox_random123_known_answers(orc):
    # Random123 kat_vectors, philox4x32-10
    assert [hex(v) for v in orc.philox4x32([0, 0, 0, 0], [0, 0])] == \
        ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c", "0x9b00dbd8"]
    assert [hex(v) for v in orc.philox4x32([0xffffffff] * 4, [0xffffffff] * 2)] == \
        ["0x408f276d", "0x41c83b0e", "0xa20bc7c6", "0x6d5451fd"]
    assert [hex(v) for v in orc.philox4x32([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                                           [0xa4093822, 0x299f31d0])] == \
        ["0xd16cfe09", "0x94fdcceb", "0x5001e420", "0x24126ea1"]


def test_uniform_open_interval_and_moments(orc):
    u = np.array([orc.uniform(7, i, 3) for i in range(20000)])
    assert u.min() > 0.0 and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 0.01
    assert abs(u.var() - 1.0 / 12.0) < 0.005
    # keyed: a different column or seed gives a different stream
    assert orc.uniform(7, 0, 3) != orc.uniform(7, 0, 4) != orc.uniform(8, 0, 3)


def test_normal_moments(orc):
    z = np.array([orc.normal(42, i) for i in range(40000)])
    assert abs(z.mean()) < 0.03
    assert abs(z.var() - 1.0) < 0.04
    assert abs(np.mean(z ** 4) - 3.0) < 0.15       # Box-Muller gives Gaussian kurtosis


# ------------------------------------------------------------------ input path
@pytest.mark.parametrize("name", ["path4.mtx", "star5.mtx", "triangle.mtx", "grid3x3.mtx", "two_components.mtx"])
def test_laplacian_matches_golden(orc, golden, name):
    a = orc.laplacian(*read_mtx_py(golden_path(name)))
    np.testing.assert_allclose(a.dense(), np.array(golden[name]), rtol=0, atol=1e-14)
    # columns sorted by row
    for j in range(a.n):
        seg = a.ind[a.ptr[j]:a.ptr[j + 1]]
        assert np.all(np.diff(seg) >= 0)


def test_dependency_counts_are_lower_neighbours(orc):
    a = orc.laplacian(*read_mtx_py(golden_path("grid3x3.mtx")))
    dep = orc.dependency_counts(a)
    d = a.dense()
    for k in range(a.n):
        assert dep[k] == np.count_nonzero(d[:k, k])


def test_explicit_zero_offdiagonal_dropped(orc):
    a = orc.laplacian(3, np.array([0, 1, 1, 2]), np.array([1, 0, 2, 1]), np.array([-1.0, -1.0, 0.0, 0.0]))
    assert a.nnz == 2 + 3          # two off-diagonals + appended diagonal


def test_ground_append_matches_golden(orc, golden):
    a = orc.laplacian(*read_mtx_py(golden_path("sddm3.mtx")))
    np.testing.assert_allclose(a.dense(), np.array(golden["sddm3.mtx"]), atol=1e-14)
    g, appended = orc.ground_append(a)
    assert appended and g.n == 4
    gd = g.dense()
    np.testing.assert_allclose(gd, np.array(golden["sddm3.mtx:grounded"]), atol=1e-14)
    np.testing.assert_allclose(gd.sum(axis=0), 0.0, atol=1e-14)
    np.testing.assert_allclose(gd, gd.T)


def test_ground_append_rejects_nondominant(orc):
    a = orc.laplacian(*read_mtx_py(golden_path("nondominant.mtx")))
    with pytest.raises(ValueError):
        orc.ground_append(a)


def test_ground_append_noop_on_laplacian(orc):
    a = orc.laplacian(*read_mtx_py(golden_path("grid3x3.mtx")))
    g, appended = orc.ground_append(a)
    assert not appended and g.n == a.n


def test_trim_last_inverts_ground_append(orc):
    a = orc.laplacian(*read_mtx_py(golden_path("sddm3.mtx")))
    g, _ = orc.ground_append(a)
    t = orc.trim_last(g)
    np.testing.assert_array_equal(t.dense(), a.dense())


def test_random_perm_valid_and_seeded(orc):
    p = orc.random_perm(1000, 5)
    assert sorted(p.tolist()) == list(range(1000))
    assert not np.array_equal(p, orc.random_perm(1000, 6))
    np.testing.assert_array_equal(p, orc.random_perm(1000, 5))
    # Fisher-Yates is unbiased: position of element 0 roughly uniform
    pos = [int(np.where(orc.random_perm(10, s) == 0)[0][0]) for s in range(3000)]
    counts = np.bincount(pos, minlength=10)
    assert counts.min() > 220 and counts.max() < 380


def test_nnz_sort_is_stable_degree_sort_of_random_perm(orc, wl):
    n, r, c, v = wl.grid_laplacian((9, 7))
    a = orc.laplacian(n, r, c, v)
    p = orc.nnz_sort_perm(a, 3)
    assert sorted(p.tolist()) == list(range(n))
    cnt = np.diff(a.ptr)
    assert np.all(np.diff(cnt[p]) >= 0)
    p1 = orc.random_perm(n, 3)
    expect = p1[np.argsort(cnt[p1], kind="stable")]
    np.testing.assert_array_equal(p, expect)


def test_permute_is_symmetric_permutation(orc, wl):
    n, r, c, v = wl.grid_laplacian((5, 4), "loguniform", seed=3)
    a = orc.laplacian(n, r, c, v)
    p = orc.random_perm(n, 11)
    b = orc.permute(a, p)
    np.testing.assert_array_equal(b.dense(), a.dense()[np.ix_(p, p)])


# ------------------------------------------------------------------ elimination
def _ldlt(f):
    L = f.dense_L()
    return L @ np.diag(f.diag) @ L.T


def _edges_lap(n, edges):
    u = np.array([e[0] for e in edges]); v = np.array([e[1] for e in edges]); w = np.array([e[2] for e in edges])
    return n, np.concatenate([u, v]).astype(np.int32), np.concatenate([v, u]).astype(np.int32), -np.concatenate([w, w])


def _exact_cases():
    rng = np.random.default_rng(0)
    path = _edges_lap(8, [(i, i + 1, float(rng.uniform(0.1, 5))) for i in range(7)])
    cycle = _edges_lap(9, [(i, (i + 1) % 9, float(rng.uniform(0.1, 5))) for i in range(9)])
    # binary tree numbered leaves first (post-order): every vertex has <= 1 later neighbour
    tree = _edges_lap(7, [(0, 2, 1.5), (1, 2, 0.5), (3, 5, 2.0), (4, 5, 3.0), (2, 6, 1.0), (5, 6, 4.0)])
    star_last = _edges_lap(6, [(i, 5, float(i + 1)) for i in range(5)])
    return {"path": path, "cycle": cycle, "tree": tree, "star_centre_last": star_last}


@pytest.mark.parametrize("case", ["path", "cycle", "tree", "star_centre_last"])
def test_factor_exact_when_every_degree_le_2(orc, case):
    """d_k <= 2 for every column makes the sampled elimination exact."""
    n, r, c, v = _exact_cases()[case]
    a = orc.laplacian(n, r, c, v)
    for seed in (1, 2, 99):
        f = orc.factor(a, seed=seed)
        assert f.stats["max_degree"] <= 2
        np.testing.assert_allclose(_ldlt(f), a.dense(), atol=1e-12 * np.abs(a.dense()).max())


def test_factor_pivot_mutation_caught(orc):
    n, r, c, v = _exact_cases()["cycle"]
    a = orc.laplacian(n, r, c, v)
    f = orc.factor(a, seed=1, mutation=orc.MUT_PIVOT)
    assert not np.allclose(_ldlt(f), a.dense(), atol=1e-8)


def _check_invariants(orc, a, f, ncomp):
    n = a.n
    low = a.ind > np.repeat(np.arange(n), np.diff(a.ptr))
    nnz_low = int(low.sum())
    deg = np.diff(f.ptr)
    for k in range(n):
        seg = f.ind[f.ptr[k]:f.ptr[k + 1]]
        assert np.all(seg > k) and np.all(np.diff(seg) > 0)
        vals = f.val[f.ptr[k]:f.ptr[k + 1]]
        if len(seg):
            assert np.all(vals < 0)
            assert abs(vals.sum() + 1.0) < 1e-12          # sum_i w_i / S = 1
            assert f.diag[k] > 0
            assert np.all(f.level[seg] > f.level[k])
        else:
            assert f.diag[k] == 0.0
    assert f.stats["zero_pivots"] == ncomp
    assert f.stats["fills"] == int(np.maximum(deg - 1, 0).sum())
    assert f.stats["gathered"] == nnz_low + f.stats["fills"]
    assert f.stats["nnz_off"] == len(f.ind)
    assert f.stats["nlevels"] == f.level.max() + 1
    # sampled Schur complements are Laplacians: L D L^T has zero row sums
    if n <= 400:
        np.testing.assert_allclose(_ldlt(f).sum(axis=1), 0.0, atol=1e-9 * np.abs(a.val).max())


@pytest.mark.parametrize("kind", ["grid2d", "grid3d", "rgg", "two_comp"])
def test_factor_invariants(orc, wl, kind):
    if kind == "grid2d":
        n, r, c, v = wl.grid_laplacian((14, 13), "loguniform", seed=1)
        ncomp = 1
    elif kind == "grid3d":
        n, r, c, v = wl.grid_laplacian((6, 6, 6))
        ncomp = 1
    elif kind == "rgg":
        n, r, c, v = wl.random_geometric_graph(350, seed=4)
        import scipy.sparse as sp
        from scipy.sparse.csgraph import connected_components
        ncomp = connected_components(sp.coo_matrix((np.ones(len(r)), (r, c)), shape=(n, n)))[0]
    else:
        n, r, c, v = read_mtx_py(golden_path("two_components.mtx"))
        ncomp = 2
    a0 = orc.laplacian(n, r, c, v)
    a = orc.permute(a0, orc.nnz_sort_perm(a0, 1))
    f = orc.factor(a, seed=5)
    _check_invariants(orc, a, f, ncomp)


def _star(weights):
    d = len(weights)
    return _edges_lap(d + 1, [(0, i + 1, float(w)) for i, w in enumerate(weights)])


def _mean_ldlt(orc, a, nseeds, mutation=0):
    acc = np.zeros((a.n, a.n))
    acc2 = np.zeros((a.n, a.n))
    for s in range(nseeds):
        m = _ldlt(orc.factor(a, seed=1000 + s, mutation=mutation))
        acc += m
        acc2 += m * m
    mean = acc / nseeds
    sd = np.sqrt(np.maximum(acc2 / nseeds - mean * mean, 0.0))
    return mean, sd


def test_factor_unbiased_on_star(orc):
    """E[L D L^T] = A: the sampled clique has expectation w_i w_j / S."""
    n, r, c, v = _star([0.5, 1.0, 2.0, 3.0, 4.5, 7.0])
    a = orc.laplacian(n, r, c, v)
    N = 3000
    mean, sd = _mean_ldlt(orc, a, N)
    tol = 5.0 * sd / np.sqrt(N) + 1e-12
    assert np.all(np.abs(mean - a.dense()) <= tol)


@pytest.mark.parametrize("mutation", [1, 2])   # MUT_FILL_WEIGHT, MUT_SAMPLE_SHIFT
def test_factor_unbiasedness_catches_sampler_mistakes(orc, mutation):
    n, r, c, v = _star([0.5, 1.0, 2.0, 3.0, 4.5, 7.0])
    a = orc.laplacian(n, r, c, v)
    N = 3000
    mean, sd = _mean_ldlt(orc, a, N, mutation=mutation)
    tol = 5.0 * sd / np.sqrt(N) + 1e-12
    assert np.any(np.abs(mean - a.dense()) > tol)


def test_factor_merges_duplicate_rows(orc):
    """Column 0 of a triangle: its fill lands on an existing edge and the two
    entries of row 2 in column 1 merge (weight summed, multiplicity 2)."""
    n, r, c, v = _edges_lap(3, [(0, 1, 1.0), (0, 2, 2.0), (1, 2, 4.0)])
    a = orc.laplacian(n, r, c, v)
    f = orc.factor(a, seed=3)
    # column 0: weights {1, 2}, S = 3, fill (2,1) weight 1*2/3 merged with 4
    assert f.diag[0] == 3.0
    assert f.diag[1] == 2.0 / 3.0 + 4.0          # merged in (row, value) order
    assert f.diag[2] == 0.0
    L = f.dense_L()
    assert L[1, 0] == -1.0 / 3.0 and L[2, 0] == -2.0 / 3.0 and L[2, 1] == -1.0
    np.testing.assert_allclose(_ldlt(f), a.dense(), atol=1e-12)     # d <= 2 everywhere: exact
    g = orc.factor(a, seed=3, mutation=orc.MUT_NO_MERGE)
    assert not np.allclose(_ldlt(g), a.dense(), atol=1e-6)


def test_factor_is_a_pure_function_of_seed(orc, wl):
    n, r, c, v = wl.grid_laplacian((10, 10), "loguniform", seed=2)
    a = orc.laplacian(n, r, c, v)
    f1, f2, f3 = orc.factor(a, seed=9), orc.factor(a, seed=9), orc.factor(a, seed=10)
    np.testing.assert_array_equal(f1.ind, f2.ind)
    np.testing.assert_array_equal(f1.val, f2.val)
    assert len(f1.ind) != len(f3.ind) or not np.array_equal(f1.val, f3.val)


def test_factor_trim_last_drops_ground(orc, wl):
    n, r, c, v = wl.grid_sddm((4, 4, 3))
    a = orc.laplacian(n, r, c, v)
    g, appended = orc.ground_append(a)
    assert appended
    f = orc.factor(g, seed=1)
    t = orc.factor_trim_last(f)
    assert t.n == n
    assert np.all(t.ind < n)
    full = f.dense_L()[:n, :n]
    np.testing.assert_array_equal(t.dense_L(), full)


# ------------------------------------------------------------------ preconditioner / PCG
def _dense_precond(f, r):
    L = f.dense_L()
    y = np.linalg.solve(L, r)
    d = f.diag
    y = np.where(d != 0.0, y / np.where(d != 0.0, d, 1.0), y)      # solver.hpp:64-84
    return np.linalg.solve(L.T, y)


def test_precond_matches_dense_triangular_solves(orc, wl):
    n, r, c, v = wl.grid_laplacian((12, 11), "loguniform", seed=5)
    a0 = orc.laplacian(n, r, c, v)
    a = orc.permute(a0, orc.nnz_sort_perm(a0, 2))
    f = orc.factor(a, seed=4)
    rr = np.random.default_rng(1).standard_normal(n)
    z = orc.precond(f, rr)
    zd = _dense_precond(f, rr)
    np.testing.assert_allclose(z, zd, rtol=1e-9, atol=1e-9 * np.abs(zd).max())
    z32 = orc.precond(f, rr, precision=32)
    np.testing.assert_allclose(z32, zd, rtol=0, atol=2e-3 * np.abs(zd).max())
    zbad = orc.precond(f, rr, mutation=orc.MUT_PRECOND_SIGN)
    assert not np.allclose(zbad, zd, atol=1e-6 * np.abs(zd).max())


def test_pcg_with_exact_factor_converges_in_one_iteration(orc):
    n, r, c, v = _exact_cases()["path"]
    a = orc.laplacian(n, r, c, v)
    f = orc.factor(a, seed=1)
    b = np.random.default_rng(2).standard_normal(n); b -= b.mean()
    x, st = orc.pcg(a, f, b, tol=1e-10)
    assert st["iterations"] == 1 and st["rc"] == 0
    assert st["rel_res_true"] < 1e-12


def test_pcg_zero_rhs_takes_no_iteration(orc, wl):
    n, r, c, v = wl.grid_laplacian((6, 6))
    a = orc.laplacian(n, r, c, v)
    f = orc.factor(a, seed=1)
    x, st = orc.pcg(a, f, np.zeros(n), tol=1e-6)
    assert st["iterations"] == 0 and np.all(x == 0)


def test_pcg_matches_scipy_cg(orc, wl):
    """Same preconditioner, textbook CG (scipy): identical iteration counts on
    the true-residual rule, same solution."""
    import scipy.sparse.linalg as sla
    n, r, c, v = wl.grid_laplacian((20, 20), "loguniform", seed=6)
    a0 = orc.laplacian(n, r, c, v)
    a = orc.permute(a0, orc.nnz_sort_perm(a0, 1))
    f = orc.factor(a, seed=2)
    b = np.random.default_rng(3).standard_normal(n); b -= b.mean()
    x, st = orc.pcg(a, f, b, tol=1e-8, stop_rule=orc.STOP_TRUE_RES, maxit=500)
    A = a.scipy().tocsr()
    M = sla.LinearOperator((n, n), matvec=lambda rr: orc.precond(f, rr))
    its = [0]
    xs, info = sla.cg(A, b, rtol=1e-8, atol=0.0, M=M, maxiter=500, callback=lambda xk: its.__setitem__(0, its[0] + 1))
    assert info == 0
    assert abs(its[0] - st["iterations"]) <= 1
    # Laplacian: compare modulo the null space
    dx = (x - x.mean()) - (xs - xs.mean())
    assert np.linalg.norm(dx) <= 1e-6 * np.linalg.norm(xs - xs.mean())


def test_pcg_reference_rule_matches_recurrence(orc, wl):
    """Reference rule: stop when sqrt(r'z / r0'z0) <= tol (solver.hpp:536)."""
    n, r, c, v = wl.grid_laplacian((64, 64))
    a0 = orc.laplacian(n, r, c, v)
    a = orc.permute(a0, orc.nnz_sort_perm(a0, 1))
    f = orc.factor(a, seed=1)
    b = np.array([orc.normal(42, i) for i in range(n)]); b -= b.mean()
    x, st = orc.pcg(a, f, b, tol=1e-6)
    assert st["converged"] == 1
    assert np.sqrt(st["rz"] / st["rz0"]) <= 1e-6
    assert 10 <= st["iterations"] <= 60
    assert st["rel_res_true"] < 1e-4
