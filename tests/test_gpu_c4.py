"""The batched configuration on the GPU at reduced sizes of its own
workloads: 64 right-hand sides on a 2D grid with uniform[0.5, 2] weights,
random-geometric-graph batches of 10k-20k vertices, both in FP32, and the two
solved together in one launch sequence (rc_solve_group)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _rgg_sizes(count, lo=10_000, hi=20_000):
    return [lo + (i * (hi - lo)) // max(count - 1, 1) for i in range(count)]


def _a_norm_close(orc, sys_, x, xo, tol):
    """||x - xo||_A <= tol ||xo||_A: PCG minimises the A-norm error, so two
    runs stopped at the same tolerance agree in it even where the 2-norm of a
    badly conditioned system's low-energy modes differs."""
    d = x - xo
    return np.sqrt(max(d @ orc.spmv(sys_, d), 0.0)) <= tol * np.sqrt(xo @ orc.spmv(sys_, xo))


def _borderline(so):
    # a run stopping with sqrt(r'z / r0'z0) within 4x of the tolerance may move
    # by an iteration either way under a reordered dot product
    return so["rz0"] > 0 and np.sqrt(so["rz"] / so["rz0"]) > 0.25e-6


def _grid_solver(rc, wl, side, seed=11):
    n, r, c, v = wl.grid_laplacian((side, side), "uniform_0.5_2", seed=seed)
    s = rc.Solver(device=0, seed=1, precision=32)
    s.set_triplets(n, r, c, v)
    s.order(rc.ORDER_NNZ_SORT, 1)
    s.factor()
    return s, (n, r, c, v)


def _rhs_block(s, rc, k, seed0=42):
    return np.stack([s.make_rhs(rc.RHS_NORMAL_PROJECTED, seed0 + q) for q in range(k)], axis=1)


def test_64_rhs_on_uniform_grid_match_oracle(rc, orc, wl):
    say("[gpu] c4 64 rhs 256^2 fp32")
    s, (n, r, c, v) = _grid_solver(rc, wl, 256)
    rp, ci, va, dg = s.get_factor()
    B = _rhs_block(s, rc, 64)
    s.set_rhs_multi(B)
    sts = s.solve_multi_resident(tol=1e-6)
    X = s.get_solution_multi()
    assert s.multi_stats() == sts
    fac, sys_, perm, _ = orc.prepare_system(n, r, c, v, order="nnz-sort", order_seed=1)
    f = orc.factor(fac, seed=1)
    mask = np.ones(len(ci), bool)
    mask[rp[:-1]] = False
    np.testing.assert_array_equal(ci[mask], f.ind)
    np.testing.assert_array_equal(va[mask], f.val.astype(np.float32).astype(np.float64))
    for q in range(64):
        xo, so = orc.pcg(sys_, f, B[perm, q], tol=1e-6, precision=32)
        lim = 2 if _borderline(so) else 1
        assert abs(sts[q]["iterations"] - so["iterations"]) <= lim, (q, sts[q], so)
        assert _a_norm_close(orc, sys_, X[perm, q], xo, 1e-5), q
        assert sts[q]["converged"] == 1
    s.close()


def test_rgg_batch_of_10k_20k_vertices_matches_oracle(rc, orc, wl):
    say("[gpu] c4 rgg batch 6 x 10k-20k fp32")
    systems = [wl.random_geometric_graph(m, seed=1000 + i) for i, m in enumerate(_rgg_sizes(6))]
    s = rc.Solver(device=0, seed=1, precision=32)
    s.set_batch(systems)
    s.order(rc.ORDER_NNZ_SORT, 1)
    s.factor()
    k, off = s.batch()
    b = s.make_rhs(rc.RHS_NORMAL_PROJECTED, 42)
    x, sts = s.solve_batch(b, tol=1e-6)
    # oracle: each system alone is the block-diagonal PCG with the other
    # systems' right-hand sides zeroed
    n = int(off[-1])
    R, C, V = [], [], []
    for q, (m, r, c, v) in enumerate(systems):
        R.append(np.asarray(r) + off[q]); C.append(np.asarray(c) + off[q]); V.append(np.asarray(v))
    fac, sys_, perm, _ = orc.prepare_system(n, np.concatenate(R), np.concatenate(C), np.concatenate(V),
                                            order="nnz-sort", order_seed=1)
    f = orc.factor(fac, seed=1)
    for q in range(k):
        bq = np.zeros(n)
        bq[off[q]:off[q + 1]] = b[off[q]:off[q + 1]]
        xo, so = orc.pcg(sys_, f, bq[perm], tol=1e-6, precision=32)
        xq = np.zeros(n)
        xq[off[q]:off[q + 1]] = x[off[q]:off[q + 1]]
        lim = 2 if _borderline(so) else 1
        assert abs(sts[q]["iterations"] - so["iterations"]) <= lim, (q, sts[q], so)
        assert _a_norm_close(orc, sys_, xq[perm], xo, 1e-5), q
        assert sts[q]["converged"] == 1
    s.close()


def test_group_solve_equals_separate_solves(rc, wl):
    """A 64-RHS grid and an RGG batch in one launch sequence give exactly the
    separate solves' iterates (same kernels, same arguments)."""
    say("[gpu] c4 group solve")
    g, _ = _grid_solver(rc, wl, 128)
    B = _rhs_block(g, rc, 64)
    systems = [wl.random_geometric_graph(m, seed=1000 + i) for i, m in enumerate(_rgg_sizes(4))]
    bt = rc.Solver(device=0, seed=1, precision=32)
    bt.set_batch(systems)
    bt.order(rc.ORDER_NNZ_SORT, 1)
    bt.factor()
    b = bt.make_rhs(rc.RHS_NORMAL_PROJECTED, 42)
    one, _ = _grid_solver(rc, wl, 64, seed=3)
    b1 = one.make_rhs(rc.RHS_NORMAL_PROJECTED, 7)
    # separately
    g.set_rhs_multi(B)
    st_g = g.solve_multi_resident(tol=1e-6)
    X_sep = g.get_solution_multi()
    bt.set_rhs(b)
    bt.solve_resident(1e-6)
    st_b = bt.batch_stats()
    x_sep = bt.get_solution()
    one.set_rhs(b1)
    st_1 = one.solve_resident(1e-6)
    x1_sep = one.get_solution()
    # together
    g.set_rhs_multi(B)
    bt.set_rhs(b)
    one.set_rhs(b1)
    summ = rc.solve_group([g, bt, one], tol=1e-6)
    np.testing.assert_array_equal(g.get_solution_multi(), X_sep)
    np.testing.assert_array_equal(bt.get_solution(), x_sep)
    np.testing.assert_array_equal(one.get_solution(), x1_sep)
    assert [d["iterations"] for d in g.multi_stats()] == [d["iterations"] for d in st_g]
    assert [d["iterations"] for d in bt.batch_stats()] == [d["iterations"] for d in st_b]
    assert summ[0]["iterations"] == max(d["iterations"] for d in st_g)
    assert summ[1]["iterations"] == max(d["iterations"] for d in st_b)
    assert summ[2]["iterations"] == st_1["iterations"]
    assert all(d["converged"] for d in summ)
    # a handle twice, or a member without a resident right-hand side, is refused
    with pytest.raises(rc.RcholError):
        rc.solve_group([g, g])
    fresh, _ = _grid_solver(rc, wl, 16)
    with pytest.raises(rc.RcholError):
        rc.solve_group([g, fresh])
    for s in (g, bt, one, fresh):
        s.close()


def test_multi_rhs_kernel_timing(rc, wl):
    say("[gpu] multi-rhs kernel timing")
    s, _ = _grid_solver(rc, wl, 96)
    k = 64
    B = _rhs_block(s, rc, k)
    with pytest.raises(rc.RcholError):
        s.time_kernel(rc.KERNEL_MULTI_SPMM, reps=1)          # no multi-RHS solve yet
    s.set_rhs_multi(B)
    sts = s.solve_multi_resident(tol=1e-6)
    X = s.get_solution_multi()
    n, nnz = s.size()
    rp, ci, va, dg = s.get_factor()
    nnz_l = len(ci) - n
    for kid in (rc.KERNEL_MULTI_SPMM, rc.KERNEL_MULTI_TRSV_FWD, rc.KERNEL_MULTI_TRSV_BWD):
        ms, by = s.time_kernel(kid, reps=2)
        assert ms > 0
        if kid == rc.KERNEL_MULTI_SPMM:
            assert by == 12 * nnz + 4 * (n + 1) + 16 * k * n
        else:
            assert by == 8 * nnz_l + 4 * (n + 1) + 2 * 4 * k * n
    # timing overwrote the vectors: a new solve gives the same answer again
    with pytest.raises(rc.RcholError):
        s.get_solution_multi()
    sts2 = s.solve_multi_resident(tol=1e-6)
    np.testing.assert_array_equal(s.get_solution_multi(), X)
    assert [d["iterations"] for d in sts2] == [d["iterations"] for d in sts]
    s.close()
