"""C ABI and host-side input path of librchol.so, checked on the CPU through a
host-only handle (device = -1): symbol exports, Matrix Market reading,
orderings, ground append and RHS generation against the oracle."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT, golden_path, read_mtx_py


def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "rchol.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(rc_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(rc):
    lib = rc.load_library()
    declared = _declared_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(rc.EXPORTS)


def _host(rc):
    return rc.Solver(device=-1)


def _csc_dense(n, ptr, ind, val):
    a = np.zeros((n, n))
    for j in range(n):
        for p in range(ptr[j], ptr[j + 1]):
            a[ind[p], j] += val[p]
    return a


@pytest.mark.parametrize("name", ["path4.mtx", "star5.mtx", "triangle.mtx", "grid3x3.mtx", "two_components.mtx"])
def test_mtx_reader_and_laplacian(rc, orc, golden, name):
    s = _host(rc)
    s.load_mtx(golden_path(name))
    n, ptr, ind, val = s.prepared_csc()
    np.testing.assert_array_equal(_csc_dense(n, ptr, ind, val), np.array(golden[name]))
    a = orc.laplacian(*read_mtx_py(golden_path(name)))
    np.testing.assert_array_equal(ptr, a.ptr)
    np.testing.assert_array_equal(ind, a.ind)
    np.testing.assert_array_equal(val, a.val)


@pytest.mark.parametrize("method", ["nnz", "random"])
def test_orderings_match_oracle(rc, orc, wl, method):
    n, r, c, v = wl.grid_laplacian((11, 9), "loguniform", seed=4)
    s = _host(rc)
    s.set_triplets(n, r, c, v)
    s.order(rc.ORDER_NNZ_SORT if method == "nnz" else rc.ORDER_RANDOM, 17)
    p = s.permutation()
    a = orc.laplacian(n, r, c, v)
    expect = orc.nnz_sort_perm(a, 17) if method == "nnz" else orc.random_perm(n, 17)
    np.testing.assert_array_equal(p, expect)
    _, ptr, ind, val = s.prepared_csc()
    b = orc.permute(a, expect)
    np.testing.assert_array_equal(ptr, b.ptr)
    np.testing.assert_array_equal(ind, b.ind)
    np.testing.assert_array_equal(val, b.val)


def test_sddm_order_then_ground_append(rc, orc, wl):
    n, r, c, v = wl.grid_sddm((5, 4, 3))
    s = _host(rc)
    s.set_triplets(n, r, c, v, kind=rc.INPUT_SDDM)
    s.order(rc.ORDER_NNZ_SORT, 1)
    nf, ptr, ind, val = s.prepared_csc()
    fac, sys, perm, appended = orc.prepare_system(n, r, c, v, kind="sddm", order="nnz-sort", order_seed=1)
    assert appended and nf == n + 1
    np.testing.assert_array_equal(ptr, fac.ptr)
    np.testing.assert_array_equal(ind, fac.ind)
    np.testing.assert_array_equal(val, fac.val)
    assert s.size()[0] == n


def test_grounded_file_round_trip(rc, tmp_path, wl):
    n, r, c, v = wl.grid_sddm((4, 4))
    s = _host(rc)
    s.set_triplets(n, r, c, v, kind=rc.INPUT_SDDM)
    s.order(rc.ORDER_NNZ_SORT, 3)
    path = tmp_path / "g-nnz-sorted.mtx"
    s.write_prepared_mtx(path)
    _, ptr, ind, val = s.prepared_csc()
    t = _host(rc)
    t.load_mtx(path, kind=rc.INPUT_GROUNDED)
    assert t.size()[0] == n
    _, ptr2, ind2, val2 = t.prepared_csc()
    np.testing.assert_array_equal(ptr, ptr2)
    np.testing.assert_array_equal(ind, ind2)
    np.testing.assert_array_equal(val, val2)


def test_rhs_generators(rc, orc):
    s = _host(rc)
    s.load_mtx(golden_path("two_components.mtx"))
    b = s.make_rhs(rc.RHS_NORMAL, 42)
    np.testing.assert_array_equal(b, [orc.normal(42, i) for i in range(5)])
    bp = s.make_rhs(rc.RHS_NORMAL_PROJECTED, 42)
    assert abs(bp[:2].sum()) < 1e-15 and abs(bp[2:].sum()) < 1e-14
    bu = s.make_rhs(rc.RHS_ZERO_SUM_UNIFORM, 0)
    assert abs(bu.sum()) < 1e-14 and np.all(np.abs(bu) < 1.0)


def test_reference_rhs_matches_mt19937(rc):
    """RHS_ZERO_SUM_UNIFORM = solver.hpp:43-61: std::mt19937(seed) through
    uniform_real_distribution<double>(0,1) (two 32-bit draws per value,
    low word first), minus the mean.  numpy's legacy MT19937 seeded with an
    integer uses the same init_genrand, so its raw words pin the stream."""
    s = _host(rc)
    s.load_mtx(golden_path("grid3x3.mtx"))
    b = s.make_rhs(rc.RHS_ZERO_SUM_UNIFORM, 0)
    rs = np.random.RandomState(0)
    words = rs.randint(0, 2 ** 32, size=18, dtype=np.uint64)   # raw 32-bit outputs
    u = (words[0::2].astype(np.float64) + words[1::2].astype(np.float64) * 2.0 ** 32) / 2.0 ** 64
    np.testing.assert_allclose(b, u - u.mean(), rtol=0, atol=1e-15)


@pytest.mark.parametrize("name,code", [("nonsquare.mtx", -3), ("truncated.mtx", -3), ("badheader.mtx", -3),
                                       ("missing.mtx", -2)])
def test_malformed_inputs_rejected(rc, name, code):
    s = _host(rc)
    with pytest.raises(rc.RcholError) as e:
        s.load_mtx(golden_path(name))
    assert e.value.status == code


def test_nondominant_sddm_rejected(rc):
    s = _host(rc)
    s.load_mtx(golden_path("nondominant.mtx"), kind=rc.INPUT_SDDM)
    with pytest.raises(rc.RcholError) as e:
        s.prepared_csc()
    assert e.value.status == rc.ERR_NOT_DOMINANT


def test_nonsymmetric_rejected(rc):
    s = _host(rc)
    with pytest.raises(rc.RcholError) as e:
        s.set_triplets(3, [0, 1], [1, 2], [-1.0, -1.0])
    assert e.value.status == rc.ERR_NOT_SYMMETRIC


def test_device_calls_on_host_handle_fail_cleanly(rc):
    s = _host(rc)
    s.load_mtx(golden_path("grid3x3.mtx"))
    with pytest.raises(rc.RcholError) as e:
        s.factor()
    assert e.value.status == rc.ERR_NO_DEVICE
    with pytest.raises(rc.RcholError) as e:
        s.solve(np.zeros(9))
    assert e.value.status in (rc.ERR_NO_DEVICE, rc.ERR_STATE)


def test_out_of_order_calls(rc):
    s = _host(rc)
    with pytest.raises(rc.RcholError) as e:
        s.order(rc.ORDER_NNZ_SORT, 1)
    assert e.value.status == rc.ERR_STATE


# ------------------------------------------------------------------ AMD ordering
def _symbolic_fill(n, ptr, ind):
    """nnz of the strict lower Cholesky factor of the pattern (dense symbolic)."""
    M = np.zeros((n, n), dtype=bool)
    for j in range(n):
        M[ind[ptr[j]:ptr[j + 1]], j] = True
    cnt = 0
    for k in range(n):
        nb = np.where(M[k + 1:, k])[0] + k + 1
        cnt += len(nb)
        if len(nb):
            M[np.ix_(nb, nb)] = True
    return cnt


def _exact_min_degree_fill(n, edges):
    """Fill of exact minimum degree (ties by smallest index) by brute force."""
    adj = [set() for _ in range(n)]
    for u, v in edges:
        adj[u].add(v); adj[v].add(u)
    alive = set(range(n))
    fill = 0
    while alive:
        p = min(alive, key=lambda v: (len(adj[v] & alive), v))
        nb = adj[p] & alive
        fill += len(nb)
        for a in nb:
            adj[a] |= nb - {a}
        alive.remove(p)
    return fill


@pytest.mark.parametrize("dims", [(12, 12), (7, 6, 5), (30, 9)])
def test_amd_valid_and_reduces_fill(rc, wl, dims):
    n, r, c, v = wl.grid_laplacian(dims)
    s = _host(rc)
    s.set_triplets(n, r, c, v)
    nat = _symbolic_fill(*s.prepared_csc()[:3])
    s.order(rc.ORDER_AMD, 1)
    p = s.permutation()
    assert sorted(p.tolist()) == list(range(n))
    amd = _symbolic_fill(*s.prepared_csc()[:3])
    assert amd < 0.75 * nat
    # within 30 % of SuperLU's multiple minimum degree on A + A^T
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    A = sp.coo_matrix((v, (r, c)), shape=(n, n)).tocsc() + 10.0 * sp.eye(n, format="csc")
    perm_c = sla.splu(A, permc_spec="MMD_AT_PLUS_A").perm_c
    B = sp.csc_matrix(A[perm_c][:, perm_c])
    B.sort_indices()
    mmd = _symbolic_fill(n, B.indptr, B.indices)
    assert amd <= 1.3 * mmd


def test_amd_close_to_exact_minimum_degree_on_small_graphs(rc):
    rng = np.random.default_rng(3)
    for trial in range(25):
        n = int(rng.integers(5, 10))
        edges = set()
        for _ in range(int(rng.integers(n, 3 * n))):
            u, w = rng.integers(0, n, 2)
            if u != w:
                edges.add((min(u, w), max(u, w)))
        edges = sorted(edges)
        if not edges:
            continue
        u = np.array([e[0] for e in edges]); w_ = np.array([e[1] for e in edges])
        s = _host(rc)
        s.set_triplets(n, np.concatenate([u, w_]), np.concatenate([w_, u]), -np.ones(2 * len(edges)))
        s.order(rc.ORDER_AMD, 1)
        amd = _symbolic_fill(*s.prepared_csc()[:3])
        md = _exact_min_degree_fill(n, edges)
        assert amd <= md + 2, (n, edges, amd, md)


def test_batch_input_host(rc, wl):
    """rc_set_batch: block-diagonal assembly, offsets, per-system RHS
    projection, and the error paths (no device needed)."""
    s = rc.Solver(device=-1)
    sysA = wl.grid_laplacian((4, 5))
    sysB = wl.grid_laplacian((3, 3, 2), "uniform_0.5_2", seed=2)
    s.set_batch([sysA, sysB])
    k, off = s.batch()
    assert k == 2 and list(off) == [0, 20, 38]
    n, nnz = s.size()
    assert n == 38
    b = s.make_rhs(rc.RHS_NORMAL_PROJECTED, 42)
    assert abs(b[:20].sum()) < 1e-12 and abs(b[20:].sum()) < 1e-12
    # the prepared Laplacian is block diagonal
    s.order(rc.ORDER_AS_IS, 1)
    nn, ptr, ind, val = s.prepared_csc()
    for j in range(nn):
        rows = ind[ptr[j]:ptr[j + 1]]
        assert np.all((rows < 20) == (j < 20))
    # a plain handle reports one system
    t = rc.Solver(device=-1)
    t.set_triplets(*sysA)
    assert t.batch()[0] == 1
    # an index outside its own system is rejected
    n0, r0, c0, v0 = sysA
    bad = (n0, np.array(r0), np.array(c0), np.array(v0))
    bad[1][0] = n0
    with pytest.raises(rc.RcholError):
        s.set_batch([bad, sysB])
    with pytest.raises(rc.RcholError):
        s.set_batch([(0, np.array([], np.int32), np.array([], np.int32), np.array([]))])
