"""Handle state across systems on the GPU: pools and multi-RHS state are
sized for the system they were made for, and convergence flags follow the
selected stop rule."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_multi_rhs_state_does_not_survive_a_new_matrix(rc, wl):
    s = rc.Solver(device=0, seed=3)
    n, r, c, v = wl.grid_laplacian((8, 8))
    s.set_triplets(n, r, c, v)
    s.order(rc.ORDER_NNZ_SORT, 1)
    s.factor()
    B = np.stack([s.make_rhs(rc.RHS_NORMAL_PROJECTED, 40 + q) for q in range(3)], axis=1)
    s.set_rhs_multi(B)
    s.solve_multi_resident(tol=1e-6)
    # a larger matrix on the same handle: the right-hand sides and solutions
    # held for the old one must not be used
    n2, r2, c2, v2 = wl.grid_laplacian((20, 20))
    s.set_triplets(n2, r2, c2, v2)
    s.order(rc.ORDER_NNZ_SORT, 1)
    s.factor()
    for call in (lambda: s.solve_multi_resident(tol=1e-6), s.get_solution_multi):
        with pytest.raises(rc.RcholError) as e:
            call()
        assert e.value.status == rc.ERR_STATE
    # after new right-hand sides it works again
    B2 = np.stack([s.make_rhs(rc.RHS_NORMAL_PROJECTED, 50 + q) for q in range(3)], axis=1)
    s.set_rhs_multi(B2)
    st = s.solve_multi_resident(tol=1e-6)
    assert all(d["converged"] for d in st)
    assert s.get_solution_multi().shape == (n2, 3)
    s.close()


def test_pools_sized_for_each_new_system(rc, wl):
    """A tiny system, then a much larger one on the same handle: no pool
    regrowth, and the same factor as a fresh handle."""
    s = rc.Solver(device=0, seed=7)
    n, r, c, v = wl.grid_laplacian((4, 4))
    s.set_triplets(n, r, c, v)
    s.order(rc.ORDER_NNZ_SORT, 1)
    s.factor()
    n2, r2, c2, v2 = wl.grid_laplacian((40, 40, 20), "loguniform", seed=2)
    s.set_triplets(n2, r2, c2, v2)
    s.order(rc.ORDER_NNZ_SORT, 1)
    st = s.factor()
    assert st["pool_regrowths"] == 0
    got = s.get_factor()
    t = rc.Solver(device=0, seed=7)
    t.set_triplets(n2, r2, c2, v2)
    t.order(rc.ORDER_NNZ_SORT, 1)
    t.factor()
    want = t.get_factor()
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)
    s.close()
    t.close()


def test_multi_rhs_true_residual_rule_reports_unconverged(rc, wl):
    s = rc.Solver(device=0, seed=3, stop_rule=rc.STOP_TRUE_RESIDUAL, maxit=3)
    n, r, c, v = wl.grid_laplacian((30, 30), "loguniform", seed=5)
    s.set_triplets(n, r, c, v)
    s.order(rc.ORDER_NNZ_SORT, 1)
    s.factor()
    B = np.stack([s.make_rhs(rc.RHS_NORMAL_PROJECTED, 60 + q) for q in range(2)], axis=1)
    s.set_rhs_multi(B)
    st = s.solve_multi_resident(tol=1e-10)
    for d in st:
        assert d["iterations"] == 3 and d["converged"] == 0
    # the single-RHS path agrees
    _, so = s.solve(B[:, 0], tol=1e-10)
    assert so["iterations"] == 3 and so["converged"] == 0
    s.close()


@pytest.mark.parametrize("name", ["grid3d_24_log", "rgg_20000", "star_3000"])
def test_queue_driven_tail_equals_round_kernels(rc, wl, monkeypatch, name):
    """The narrow rounds eliminated by one queue-driven launch give the same
    factor and the same level schedule as the round kernels."""
    print(f"[gpu] queue-driven tail {name}", flush=True)
    if name == "grid3d_24_log":
        n, r, c, v = wl.grid_laplacian((24, 24, 24), "loguniform", seed=5)
    elif name == "rgg_20000":
        n, r, c, v = wl.random_geometric_graph(20000, seed=4)
    else:
        d = 3000                                    # one column above the large-LDS capacity
        u = np.zeros(d, np.int64)
        w = np.arange(1, d + 1)
        n, r, c, v = wl.laplacian_triplets(d + 1, u, w, np.linspace(0.5, 3.0, d))
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("RCHOL_ASYNC_TAIL", mode)
        s = rc.Solver(device=0, seed=9)
        s.set_triplets(n, r, c, v)
        s.order(rc.ORDER_AS_IS if name.startswith("star") else rc.ORDER_NNZ_SORT, 1)
        st = s.factor()
        b = s.make_rhs(rc.RHS_NORMAL_PROJECTED, 3)
        x, so = s.solve(b, tol=1e-8)
        out[mode] = (st, s.get_factor(), x, so)
        s.close()
    a, b_ = out["0"], out["1"]
    for k in ("nnz_factor", "fills", "nlevels", "max_degree", "zero_pivots"):
        assert a[0][k] == b_[0][k], k
    for p, q in zip(a[1], b_[1]):
        np.testing.assert_array_equal(p, q)
    np.testing.assert_array_equal(a[2], b_[2])
    assert a[3]["iterations"] == b_[3]["iterations"]
