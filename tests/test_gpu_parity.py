"""HIP path vs the CPU oracle, through the C ABI, on an MI355X.

The factor is a pure function of (matrix, ordering, seed), so the GPU factor is
compared bit for bit with the oracle's (FP64 values exactly; FP32 values equal
to the oracle's values rounded to float).  PCG results are compared within
floating-point tolerance: the dot products reduce in a different order, so
iteration counts may differ by one and the iterates agree to ~1e-8 relative
(FP64 factor) / the FP32 factor's precision.
"""
import sys

import numpy as np
import pytest

from conftest import golden_path, read_mtx_py

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(wl, name):
    if name == "grid2d_16_log":
        return wl.grid_laplacian((16, 16), "loguniform", seed=7), "laplacian"
    if name == "grid3d_8":
        return wl.grid_laplacian((8, 8, 8)), "laplacian"
    if name == "grid3d_20_log":
        return wl.grid_laplacian((20, 20, 20), "loguniform", seed=7), "laplacian"
    if name == "grid3d_64_log":
        return wl.grid_laplacian((64, 64, 64), "loguniform", seed=7), "laplacian"
    if name == "rgg_3000":
        return wl.random_geometric_graph(3000, seed=3), "laplacian"
    if name == "sddm_7":
        return wl.grid_sddm((7, 7, 7)), "sddm"
    if name == "two_comp":
        return read_mtx_py(golden_path("two_components.mtx")), "laplacian"
    if name in ("complete_300", "er_600"):
        # dense graphs: late columns gather many fills of the same row (runs
        # of duplicates longer than the in-register merge takes) and
        # columns of 65-512 entries
        rng = np.random.default_rng(5)
        n = 300 if name == "complete_300" else 600
        iu, ju = np.triu_indices(n, 1)
        if name == "er_600":
            keep = rng.random(len(iu)) < 0.08
            iu, ju = iu[keep], ju[keep]
        w = rng.uniform(0.5, 2.0, len(iu))
        return wl.laplacian_triplets(n, iu, ju, w), "laplacian"
    if name.startswith("star_"):
        # centre eliminated first: one column with d entries.  300 uses the
        # large-LDS kernel, 3000 its global scratch, 9000 a scratch regrowth
        d = int(name.split("_")[1])
        u = np.zeros(d, np.int64); v = np.arange(1, d + 1)
        w = np.linspace(0.5, 3.0, d)
        return wl.laplacian_triplets(d + 1, u, v, w), "laplacian_asis"
    raise KeyError(name)


CASES = ["grid2d_16_log", "grid3d_8", "grid3d_20_log", "rgg_3000", "sddm_7", "two_comp", "star_300", "star_3000",
         "star_9000", "complete_300", "er_600"]


def _solver(rc, wl, name, prec=64, seed=11, **kw):
    (n, r, c, v), kind = _case(wl, name)
    s = rc.Solver(seed=seed, precision=prec, **kw)
    s.set_triplets(n, r, c, v, kind=rc.INPUT_SDDM if kind == "sddm" else rc.INPUT_LAPLACIAN)
    s.order(rc.ORDER_AS_IS if kind == "laplacian_asis" else rc.ORDER_NNZ_SORT, 1)
    return s, (n, r, c, v), kind


def _oracle(orc, data, kind, seed=11):
    n, r, c, v = data
    order = "as-is" if kind == "laplacian_asis" else "nnz-sort"
    fac, sys_, perm, appended = orc.prepare_system(n, r, c, v, kind="sddm" if kind == "sddm" else "laplacian",
                                                   order=order, order_seed=1)
    f_full = orc.factor(fac, seed=seed)
    f = orc.factor_trim_last(f_full) if appended else f_full
    return f_full, f, sys_, perm


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("name", CASES)
def test_factor_bitwise_equals_oracle(rc, orc, wl, name, prec):
    say(f"[gpu] factor parity case={name} prec={prec}")
    s, data, kind = _solver(rc, wl, name, prec)
    st = s.factor()
    rp, ci, va, dg = s.get_factor()
    f_full, f, _, _ = _oracle(orc, data, kind)
    n = f.n
    assert len(rp) == n + 1
    # unit diagonal first, then the oracle's off-diagonal entries
    assert np.all(ci[rp[:-1]] == np.arange(n)) and np.all(va[rp[:-1]] == 1.0)
    mask = np.ones(len(ci), bool)
    mask[rp[:-1]] = False
    np.testing.assert_array_equal(np.diff(rp) - 1, np.diff(f.ptr))
    np.testing.assert_array_equal(ci[mask], f.ind)
    if prec == 64:
        np.testing.assert_array_equal(va[mask], f.val)
        np.testing.assert_array_equal(dg, f.diag)
    else:
        np.testing.assert_array_equal(va[mask], f.val.astype(np.float32).astype(np.float64))
        np.testing.assert_array_equal(dg, f.diag.astype(np.float32).astype(np.float64))
    assert st["nnz_factor"] == f_full.stats["nnz_off"] + f_full.n
    assert st["fills"] == f_full.stats["fills"]
    assert st["max_degree"] == f_full.stats["max_degree"]
    assert st["zero_pivots"] == f_full.stats["zero_pivots"]
    assert st["nlevels"] == int(f.level.max()) + 1


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("name", ["grid2d_16_log", "grid3d_20_log", "rgg_3000", "sddm_7", "two_comp"])
def test_pcg_matches_oracle(rc, orc, wl, name, prec):
    say(f"[gpu] pcg parity case={name} prec={prec}")
    s, data, kind = _solver(rc, wl, name, prec)
    s.factor()
    b = s.make_rhs(rc.RHS_NORMAL if kind == "sddm" else rc.RHS_NORMAL_PROJECTED, 42)
    x, st = s.solve(b, tol=1e-6)
    _, f, sys_, perm = _oracle(orc, data, kind)
    xo, so = orc.pcg(sys_, f, b[perm], tol=1e-6, precision=prec)
    assert so["rc"] == 0
    assert abs(st["iterations"] - so["iterations"]) <= 1, (st, so)
    xg = x[perm]
    tol = 1e-7 if prec == 64 else 1e-4
    assert np.linalg.norm(xg - xo) <= tol * np.linalg.norm(xo)
    assert st["rel_res_true"] <= max(2.0 * so["rel_res_true"], 1e-9)


def test_c1_config_iterations_and_residual(rc, orc, wl):
    """64x64 unit grid, nnz-sort seed 1, FP64, N(0,1) seed 42 projected, tol 1e-6."""
    say("[gpu] C1 64x64")
    n, r, c, v = wl.grid_laplacian((64, 64))
    s = rc.Solver(seed=1, precision=64)
    s.set_triplets(n, r, c, v)
    s.order(rc.ORDER_NNZ_SORT, 1)
    st_f = s.factor()
    b = s.make_rhs(rc.RHS_NORMAL_PROJECTED, 42)
    x, st = s.solve(b, tol=1e-6)
    fac, sys_, perm, _ = orc.prepare_system(n, r, c, v)
    f = orc.factor(fac, seed=1)
    xo, so = orc.pcg(sys_, f, b[perm], tol=1e-6)
    assert st_f["nnz_factor"] == f.stats["nnz_off"] + n
    assert abs(st["iterations"] - so["iterations"]) <= 1
    assert st["converged"] == 1
    assert st["rel_res_true"] < 1e-5


def test_schedule_independence(rc, wl):
    """Any workgroup count gives the same factor (work-stealing queue)."""
    say("[gpu] schedule independence")
    outs = []
    for nb in (1, 3, 64, 2048):
        s, _, _ = _solver(rc, wl, "grid3d_8", 64, num_blocks=nb)
        s.factor()
        outs.append(s.get_factor())
        s.close()
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("cap", [1, 7])
def test_ready_list_overflow_gives_same_factor(rc, wl, cap, monkeypatch):
    """Ready-list segments shrunk to `cap` entries: nearly every append of a
    round spills into the overflow region; the factor is unchanged."""
    say(f"[gpu] ready-list overflow cap={cap}")
    s1, _, _ = _solver(rc, wl, "grid3d_20_log", 64)
    s1.factor()
    monkeypatch.setenv("RCHOL_LIST_SEG_CAP", str(cap))
    s2, _, _ = _solver(rc, wl, "grid3d_20_log", 64)
    s2.factor()
    monkeypatch.delenv("RCHOL_LIST_SEG_CAP")
    for a, b in zip(s1.get_factor(), s2.get_factor()):
        np.testing.assert_array_equal(a, b)


def test_pool_regrowth_gives_same_factor(rc, wl):
    say("[gpu] pool regrowth")
    s1, _, _ = _solver(rc, wl, "grid3d_64_log", 64)
    s1.factor()
    # one workgroup keeps the per-wave chunk slack of the pools small (the
    # deferred-column grid adds about a million entries), so a tiny pool
    # factor must overflow on 2.2M fills and regrow
    s2, _, _ = _solver(rc, wl, "grid3d_64_log", 64, pool_factor=0.05, num_blocks=1)
    st2 = s2.factor()
    assert st2["pool_regrowths"] >= 1
    for a, b in zip(s1.get_factor(), s2.get_factor()):
        np.testing.assert_array_equal(a, b)


def test_dump_and_reload_factor(rc, wl, tmp_path):
    say("[gpu] dump/load")
    s, data, kind = _solver(rc, wl, "sddm_7", 64)
    s.factor()
    b = s.make_rhs(rc.RHS_NORMAL, 42)
    x1, st1 = s.solve(b, tol=1e-8)
    prefix = str(tmp_path / "run_")
    s.dump_factor(prefix)
    lines = open(prefix + "l_sol.mtx").read().splitlines()
    assert lines[0].startswith("%%MatrixMarket matrix coordinate real general")
    t, _, _ = _solver(rc, wl, "sddm_7", 64)
    t.load_factor(prefix)
    x2, st2 = t.solve(b, tol=1e-8)
    assert st1["iterations"] == st2["iterations"]
    np.testing.assert_allclose(x1, x2, rtol=0, atol=1e-12 * np.abs(x1).max())


def test_resident_solve_and_kernel_timing(rc, wl):
    say("[gpu] resident solve + timing")
    s, _, kind = _solver(rc, wl, "grid3d_20_log", 32)
    s.factor()
    b = s.make_rhs(rc.RHS_NORMAL_PROJECTED, 42)
    x, st = s.solve(b, tol=1e-6)
    s.set_rhs(b)
    st2 = s.solve_resident(1e-6)
    assert st2["iterations"] == st["iterations"]
    np.testing.assert_array_equal(s.get_solution(), x)
    for k in (rc.KERNEL_SPMV, rc.KERNEL_TRSV_FWD, rc.KERNEL_TRSV_BWD, rc.KERNEL_PCG_ITER):
        ms, by = s.time_kernel(k, reps=3)
        assert ms > 0 and by > 0


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("tail", ["0", "64", "1024"])
def test_tail_width_limits_agree_with_oracle(rc, orc, wl, tail, prec, monkeypatch):
    """Every tail width limit (no tail; narrow; up to 1024-row levels in the
    LDS workgroup) solves the same system to the oracle's answer, with the
    factor in either precision (the tail's level count is then rarely a
    multiple of its pipeline depth, so the padding levels are exercised)."""
    say(f"[gpu] tail limit {tail} prec={prec}")
    monkeypatch.setenv("RCHOL_TAIL_ROWS", tail)
    s, data, kind = _solver(rc, wl, "grid3d_20_log", prec)
    s.factor()
    b = s.make_rhs(rc.RHS_NORMAL_PROJECTED, 42)
    x, st = s.solve(b, tol=1e-6)
    _, f, sys_, perm = _oracle(orc, data, kind)
    xo, so = orc.pcg(sys_, f, b[perm], tol=1e-6, precision=prec)
    assert abs(st["iterations"] - so["iterations"]) <= 1
    tol = 1e-7 if prec == 64 else 1e-4
    assert np.linalg.norm(x[perm] - xo) <= tol * np.linalg.norm(xo)


@pytest.mark.parametrize("order", ["length", "bucket", "groups"])
@pytest.mark.parametrize("prec", [64, 32])
def test_level_orders_agree_with_oracle(rc, orc, wl, order, prec, monkeypatch):
    """The row order inside a level (by width; width classes in input order;
    the default is input order) changes only the layout: same answer as the
    oracle."""
    say(f"[gpu] level order {order} prec={prec}")
    monkeypatch.setenv("RCHOL_LEVEL_ORDER", order)
    s, data, kind = _solver(rc, wl, "rgg_3000", prec)
    s.factor()
    b = s.make_rhs(rc.RHS_NORMAL_PROJECTED, 42)
    x, st = s.solve(b, tol=1e-6)
    _, f, sys_, perm = _oracle(orc, data, kind)
    xo, so = orc.pcg(sys_, f, b[perm], tol=1e-6, precision=prec)
    assert abs(st["iterations"] - so["iterations"]) <= 1
    tol = 1e-7 if prec == 64 else 1e-4
    assert np.linalg.norm(x[perm] - xo) <= tol * np.linalg.norm(xo)


@pytest.mark.parametrize("name", ["rgg_3000", "grid3d_20_log", "sddm_7"])
@pytest.mark.parametrize("prec", [64, 32])
def test_bisection_locality_agrees_with_oracle(rc, orc, wl, name, prec):
    """Rows of each level in the recursive bisection order of the graph
    (rc_set_locality): the factor is the oracle's bit for bit (it does not
    depend on the sweeps' layout) and PCG agrees with the oracle's."""
    say(f"[gpu] bisection locality case={name} prec={prec}")
    s, data, kind = _solver(rc, wl, name, prec)
    s.set_locality(rc.LOCALITY_BISECTION)
    s.factor()
    rp, ci, va, dg = s.get_factor()
    _, f, sys_, perm = _oracle(orc, data, kind)
    mask = np.ones(len(ci), bool)
    mask[rp[:-1]] = False
    np.testing.assert_array_equal(ci[mask], f.ind)
    if prec == 64:
        np.testing.assert_array_equal(va[mask], f.val)
    else:
        np.testing.assert_array_equal(va[mask], f.val.astype(np.float32).astype(np.float64))
    b = s.make_rhs(rc.RHS_NORMAL_PROJECTED, 42)
    x, st = s.solve(b, tol=1e-6)
    xo, so = orc.pcg(sys_, f, b[perm], tol=1e-6, precision=prec)
    assert abs(st["iterations"] - so["iterations"]) <= 1
    tol = 1e-7 if prec == 64 else 1e-4
    assert np.linalg.norm(x[perm] - xo) <= tol * np.linalg.norm(xo)


@pytest.mark.parametrize("prec", [64, 32])
def test_bisection_locality_multi_rhs_agrees_with_oracle(rc, orc, wl, prec):
    """Several right-hand sides on one factor with the bisection row order:
    every column agrees with the oracle's PCG of that right-hand side."""
    say(f"[gpu] bisection locality multi-RHS case=grid3d_20_log prec={prec}")
    s, data, kind = _solver(rc, wl, "grid3d_20_log", prec)
    s.set_locality(rc.LOCALITY_BISECTION)
    s.factor()
    _, f, sys_, perm = _oracle(orc, data, kind)
    B = np.stack([s.make_rhs(rc.RHS_NORMAL_PROJECTED, 42 + q) for q in range(3)], axis=1)
    X, sts = s.solve_multi(B, tol=1e-6)
    tol = 1e-7 if prec == 64 else 1e-4
    for q in range(3):
        xo, so = orc.pcg(sys_, f, B[perm, q], tol=1e-6, precision=prec)
        assert abs(sts[q]["iterations"] - so["iterations"]) <= 1
        assert np.linalg.norm(X[perm, q] - xo) <= tol * np.linalg.norm(xo)


@pytest.mark.parametrize("gmin", ["1", "2", "4"])
@pytest.mark.parametrize("prec", [64, 32])
def test_sweep_chunk_sizes_agree_with_oracle(rc, orc, wl, gmin, prec, monkeypatch):
    """Sweep chunks of up to 256 rows (one lane per row of a level whose rows
    have at most eight entries), 128 rows, or 64 rows (the narrowest groups):
    the same answer as the oracle.  A 2D grid has short rows in its wide
    levels, so the 256- and 128-row chunks occur."""
    say(f"[gpu] sweep chunk min group {gmin} prec={prec}")
    monkeypatch.setenv("RCHOL_SWEEP_MIN_GROUP", gmin)
    monkeypatch.setenv("RCHOL_TAIL_ROWS", "64")
    n, r, c, v = wl.grid_laplacian((120, 110), "uniform_0.5_2", seed=4)
    s = rc.Solver(device=0, seed=3, precision=prec)
    s.set_triplets(n, r, c, v)
    s.order(rc.ORDER_NNZ_SORT, 1)
    s.factor()
    b = s.make_rhs(rc.RHS_NORMAL_PROJECTED, 42)
    x, st = s.solve(b, tol=1e-6)
    fac, sys_, perm, _ = orc.prepare_system(n, r, c, v, order="nnz-sort", order_seed=1)
    f = orc.factor(fac, seed=3)
    xo, so = orc.pcg(sys_, f, b[perm], tol=1e-6, precision=prec)
    assert abs(st["iterations"] - so["iterations"]) <= 1
    tol = 1e-7 if prec == 64 else 1e-4
    assert np.linalg.norm(x[perm] - xo) <= tol * np.linalg.norm(xo)
    s.close()


def test_autotune_keeps_the_solution(rc, orc, wl):
    say("[gpu] autotune")
    s, data, kind = _solver(rc, wl, "grid3d_20_log", 32)
    s.factor()
    b = s.make_rhs(rc.RHS_NORMAL_PROJECTED, 42)
    s.set_rhs(b)
    t = s.autotune(reps=2)
    assert t["configs_tried"] > 1 and t["sweeps_ms_after"] <= t["sweeps_ms_before"]
    st = s.solve_resident(1e-6)
    x = s.get_solution()
    _, f, sys_, perm = _oracle(orc, data, kind)
    xo, so = orc.pcg(sys_, f, b[perm], tol=1e-6, precision=32)
    assert abs(st["iterations"] - so["iterations"]) <= 1
    assert np.linalg.norm(x[perm] - xo) <= 1e-4 * np.linalg.norm(xo)
    s.factor()                                   # a new factor keeps the tuned configuration
    x2, st2 = s.solve(b, tol=1e-6)
    assert abs(st2["iterations"] - so["iterations"]) <= 1


def test_sweeps_before_any_solve_on_reused_memory(rc, orc, wl):
    """The dummy slot n that ELL padding and masked lanes read is zeroed when
    the vectors are set up, so the sweeps can be timed and tuned before any
    PCG has run, also when the memory was last used by a larger solver whose
    vectors it left full of sentinels."""
    say("[gpu] sweeps before any solve, reused memory")
    big, _, _ = _solver(rc, wl, "grid3d_20_log", 32)
    big.factor()
    big.solve(big.make_rhs(rc.RHS_NORMAL_PROJECTED, 42), tol=1e-6)
    big.close()
    s, data, kind = _solver(rc, wl, "grid3d_8", 32)
    s.factor()
    b = s.make_rhs(rc.RHS_NORMAL_PROJECTED, 42)
    s.set_rhs(b)
    for k in (rc.KERNEL_TRSV_FWD, rc.KERNEL_TRSV_BWD):
        ms, by = s.time_kernel(k, reps=2)
        assert ms > 0 and by > 0
    s.autotune(reps=1)
    x, st = s.solve(b, tol=1e-6)
    _, f, sys_, perm = _oracle(orc, data, kind)
    xo, so = orc.pcg(sys_, f, b[perm], tol=1e-6, precision=32)
    assert abs(st["iterations"] - so["iterations"]) <= 1
    assert np.linalg.norm(x[perm] - xo) <= 1e-4 * np.linalg.norm(xo)


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("k", [1, 5, 64])
def test_multi_rhs_matches_oracle(rc, orc, wl, k, prec):
    """k right-hand sides in one launch sequence: every column is its own PCG."""
    say(f"[gpu] multi-rhs k={k} prec={prec}")
    s, data, kind = _solver(rc, wl, "grid3d_20_log", prec)
    s.factor()
    B = np.stack([s.make_rhs(rc.RHS_NORMAL_PROJECTED, 42 + q) for q in range(k)], axis=1)
    if k == 5:
        B[:, 3] = 0.0                      # a zero RHS stops before the loop and stays frozen
    X, sts = s.solve_multi(B, tol=1e-6)
    _, f, sys_, perm = _oracle(orc, data, kind)
    tol = 1e-7 if prec == 64 else 1e-4
    for q in range(k):
        xo, so = orc.pcg(sys_, f, B[perm, q], tol=1e-6, precision=prec)
        assert abs(sts[q]["iterations"] - so["iterations"]) <= 1, (q, sts[q], so)
        assert np.linalg.norm(X[perm, q] - xo) <= tol * max(np.linalg.norm(xo), 1e-300)
    if k == 5:
        assert sts[3]["iterations"] == 0 and np.all(X[:, 3] == 0.0)


# ---------------------------------------------------------------- batched systems
def _batch_systems(wl, many=False):
    if many:
        sizes = [20 + (37 * q) % 280 for q in range(60)]
        return [wl.random_geometric_graph(m, seed=500 + q) for q, m in enumerate(sizes)]
    return [wl.grid_laplacian((12, 12), "loguniform", seed=5),
            wl.random_geometric_graph(2500, seed=9),
            wl.grid_laplacian((6, 6, 6), "uniform_0.5_2", seed=3),
            wl.grid_laplacian((30, 20))]


def _block_diag(systems):
    off, R, C, V = 0, [], [], []
    for (n, r, c, v) in systems:
        R.append(np.asarray(r) + off)
        C.append(np.asarray(c) + off)
        V.append(np.asarray(v))
        off += n
    return off, np.concatenate(R), np.concatenate(C), np.concatenate(V)


@pytest.mark.parametrize("many,loc", [(False, False), (True, False), (False, True)])
@pytest.mark.parametrize("prec", [64, 32])
def test_batch_matches_oracle_per_system(rc, orc, wl, prec, many, loc):
    """Every system of a batch runs its own PCG: each system's iterate equals
    the oracle PCG on the block-diagonal system with the other systems'
    right-hand sides zeroed (which is that system's PCG alone); also with the
    rows of each level in every system's bisection order (loc)."""
    say(f"[gpu] batch prec={prec} many={many} loc={loc}")
    systems = _batch_systems(wl, many)
    s = rc.Solver(seed=11, precision=prec)
    s.set_batch(systems)
    if loc:
        s.set_locality(rc.LOCALITY_BISECTION)
    s.order(rc.ORDER_NNZ_SORT, 1)
    fs = s.factor()
    k, off = s.batch()
    assert k == len(systems)
    b = s.make_rhs(rc.RHS_NORMAL_PROJECTED, 42)
    zero = 2
    b[off[zero]:off[zero + 1]] = 0.0          # stops before the loop and stays at x = 0
    x, sts = s.solve_batch(b, tol=1e-6)
    n, r, c, v = _block_diag(systems)
    fac, sys_, perm, _ = orc.prepare_system(n, r, c, v, kind="laplacian", order="nnz-sort", order_seed=1)
    f = orc.factor(fac, seed=11)
    assert fs["nnz_factor"] == f.stats["nnz_off"] + n
    tol = 1e-7 if prec == 64 else 1e-4
    its = []
    for q in range(k):
        bq = np.zeros(n)
        bq[off[q]:off[q + 1]] = b[off[q]:off[q + 1]]
        xo, so = orc.pcg(sys_, f, bq[perm], tol=1e-6, precision=prec)
        xq = np.zeros(n)
        xq[off[q]:off[q + 1]] = x[off[q]:off[q + 1]]
        # A system whose oracle run stops with sqrt(r'z/r0'z0) within 4x of the
        # tolerance is borderline: a reordered dot product can move its stop
        # by an iteration either way (on the CPU, 1e-13 relative noise in b
        # alone moves such a system between 14 and 16 iterations).
        ratio = np.sqrt(so["rz"] / so["rz0"]) if so["rz0"] > 0 else 0.0
        borderline = ratio > 0.25e-6
        assert abs(sts[q]["iterations"] - so["iterations"]) <= (2 if borderline else 1), (q, sts[q], so)
        xtol = 1e-4 if borderline else tol
        assert np.linalg.norm(xq[perm] - xo) <= xtol * max(np.linalg.norm(xo), 1e-300), q
        if q != zero:
            assert sts[q]["converged"] == 1 and sts[q]["rel_res_true"] < 1e-4, (q, sts[q])
        its.append(sts[q]["iterations"])
    assert sts[zero]["iterations"] == 0 and np.all(x[off[zero]:off[zero + 1]] == 0.0)
    assert len(set(its)) > 1                   # systems stop at different iterations
    # the summary of rc_solve_resident on the same batch is the worst system
    s.set_rhs(b)
    tot = s.solve_resident(1e-6)
    assert tot["iterations"] == max(its)
    assert s.batch_stats()[1]["iterations"] == sts[1]["iterations"]


def test_multi_rhs_resident_matches_one_shot(rc, wl):
    say("[gpu] multi-rhs resident")
    s, data, kind = _solver(rc, wl, "grid3d_20_log", 32)
    s.factor()
    B = np.stack([s.make_rhs(rc.RHS_NORMAL_PROJECTED, 7 + q) for q in range(3)], axis=1)
    X1, st1 = s.solve_multi(B, tol=1e-6)
    s.set_rhs_multi(B)
    st2 = s.solve_multi_resident(1e-6)
    X2 = s.get_solution_multi()
    assert [d["iterations"] for d in st1] == [d["iterations"] for d in st2]
    np.testing.assert_array_equal(X1, X2)
    s.factor()                                      # a new factor reuses the resident right-hand sides
    st3 = s.solve_multi_resident(1e-6)
    np.testing.assert_array_equal(s.get_solution_multi(), X2)
    assert [d["iterations"] for d in st3] == [d["iterations"] for d in st2]
