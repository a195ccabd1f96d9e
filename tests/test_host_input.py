"""Host input path on the CPU (host-only handles, no device): the threaded
Matrix Market reader, header sanity, the nested-dissection ordering and the
rchol_driver command line (both reference argv forms)."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG_DIR, golden_path


def _host(rc, threads=0):
    return rc.Solver(device=-1, host_threads=threads)


def _write_mtx(path, n, rows, cols, vals, symmetric=False):
    """Matrix Market coordinate real file (1-based), written with numpy."""
    rows, cols, vals = np.asarray(rows), np.asarray(cols), np.asarray(vals)
    if symmetric:
        keep = rows >= cols
        rows, cols, vals = rows[keep], cols[keep], vals[keep]
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate real {'symmetric' if symmetric else 'general'}\n")
        f.write("% written by the test suite\n")
        f.write(f"{n} {n} {len(rows)}\n")
        np.savetxt(f, np.column_stack([rows + 1, cols + 1, vals]), fmt=["%d", "%d", "%.17g"])


@pytest.mark.parametrize("symmetric", [False, True])
def test_threaded_reader_equals_single_thread_and_triplets(rc, wl, tmp_path, symmetric):
    # a file of several MB, so the entry text is cut into many pieces
    n, r, c, v = wl.grid_laplacian((36, 36, 36), "loguniform", seed=3)
    path = str(tmp_path / "g.mtx")
    _write_mtx(path, n, r, c, v, symmetric=symmetric)
    assert os.path.getsize(path) > (1 << 20)
    ref = _host(rc)
    ref.set_triplets(n, r, c, v)
    want = ref.prepared_csc()
    base = None
    for threads in (1, 3, 8, 64):
        s = _host(rc, threads)
        s.load_mtx(path)
        got = s.prepared_csc()
        if base is None:
            base = got
            # same Laplacian as the triplets; the appended diagonal is a sum
            # whose order follows the entry order of the file
            assert got[0] == want[0]
            np.testing.assert_array_equal(got[1], want[1])
            np.testing.assert_array_equal(got[2], want[2])
            col = np.repeat(np.arange(n), np.diff(got[1]))
            off = got[2] != col
            np.testing.assert_array_equal(got[3][off], want[3][off])
            np.testing.assert_allclose(got[3][~off], want[3][~off], rtol=1e-14)
        else:
            for a, b in zip(got, base):          # file order kept: bit-identical
                np.testing.assert_array_equal(a, b)


def test_reader_keeps_only_the_declared_entries(rc, tmp_path):
    # more entry lines than the size line declares: the first nz count
    path = str(tmp_path / "extra.mtx")
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n3 3 4\n"
                "1 2 -1\n2 1 -1\n2 3 -2\n3 2 -2\n1 3 -5\n3 1 -5\n")
    s = _host(rc)
    s.load_mtx(path)
    n, ptr, ind, val = s.prepared_csc()
    assert n == 3 and ptr[-1] == 7          # path 0-1-2 with its diagonal


def test_header_claiming_huge_nz_is_rejected(rc, tmp_path):
    path = str(tmp_path / "huge.mtx")
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n3 3 1000000000000000\n1 2 -1\n2 1 -1\n")
    s = _host(rc)
    with pytest.raises(rc.RcholError) as e:
        s.load_mtx(path)
    assert e.value.status == rc.ERR_FORMAT
    assert "truncated" in s.lib.rc_last_error(s.h).decode()


def test_negative_host_threads_rejected(rc):
    with pytest.raises(rc.RcholError):
        rc.Solver(device=-1, host_threads=-1)


# ------------------------------------------------------------------ nested dissection
def _components(n, ptr, ind, alive):
    comp = -np.ones(n, np.int64)
    sizes = []
    for s0 in range(n):
        if not alive[s0] or comp[s0] >= 0:
            continue
        stack = [s0]
        comp[s0] = len(sizes)
        cnt = 0
        while stack:
            v = stack.pop()
            cnt += 1
            for w in ind[ptr[v]:ptr[v + 1]]:
                if alive[w] and comp[w] < 0:
                    comp[w] = len(sizes)
                    stack.append(w)
        sizes.append(cnt)
    return sizes


@pytest.mark.parametrize("dims", [(48, 48), (14, 14, 14), (300,)])
def test_nd_ordering_puts_a_separator_last(rc, wl, dims):
    n, r, c, v = wl.grid_laplacian(dims)
    s = _host(rc)
    s.set_triplets(n, r, c, v)
    s.order(rc.ORDER_ND, 5)
    p = s.permutation()
    assert sorted(p.tolist()) == list(range(n))
    s2 = _host(rc, 1)
    s2.set_triplets(n, r, c, v)
    s2.order(rc.ORDER_ND, 5)
    np.testing.assert_array_equal(p, s2.permutation())          # independent of the thread count
    # in the ordered graph, removing the last k vertices (the top separator)
    # splits the rest into at least two parts, none above 3/4 of the graph
    nn, ptr, ind, _ = s.prepared_csc()
    limit = int(4 * n ** (1 - 1 / len(dims))) + 2
    ok = False
    for k in range(1, limit + 1):
        alive = np.ones(nn, bool)
        alive[nn - k:] = False
        sizes = _components(nn, ptr, ind, alive)
        if len(sizes) >= 2 and max(sizes) <= 0.75 * n:
            ok = True
            break
    assert ok, "no small separator at the end of the ordering"


def test_nd_ordering_fill_and_iterations_near_nnz_sort(rc, orc, wl):
    """The oracle factors the ND-ordered grid: fill within 10 % of the
    nnz-sorted ordering's, PCG iterations within 15 %."""
    n, r, c, v = wl.grid_laplacian((24, 24, 24), "loguniform", seed=7)
    a = orc.laplacian(n, r, c, v)
    b = np.array([orc.normal(42, i) for i in range(n)])
    b -= b.mean()
    out = {}
    for name, meth in (("nnz", rc.ORDER_NNZ_SORT), ("nd", rc.ORDER_ND)):
        s = _host(rc)
        s.set_triplets(n, r, c, v)
        s.order(meth, 1)
        p = s.permutation()
        f = orc.factor(orc.permute(a, p), seed=1)
        _, st = orc.pcg(orc.permute(a, p), f, b[p], tol=1e-6)
        out[name] = (len(f.ind), st["iterations"])
    assert out["nd"][0] <= 1.10 * out["nnz"][0], out
    assert out["nd"][1] <= 1.15 * out["nnz"][1] + 1, out


# ------------------------------------------------------------------ command line
DRIVER = os.path.join(PKG_DIR, "rchol_driver")


def _dry(*args):
    if not os.path.exists(DRIVER):
        pytest.skip("rchol_driver not built")
    r = subprocess.run([DRIVER, *args, "--dry-run"], capture_output=True, text=True, timeout=60)
    return r


def test_cli_gpu_form(rc):
    r = _dry(golden_path("grid3x3.mtx"), "512", "1", "7e-7")
    assert r.returncode == 0, r.stderr
    cfg = json.loads(r.stdout.strip().splitlines()[-1])
    assert cfg["form"] == "gpu" and cfg["num_blocks"] == 512 and cfg["check_solve"] == 1
    assert cfg["tol"] == 7e-7 and cfg["kind"] == rc.INPUT_LAPLACIAN and cfg["n"] == 9 and not cfg["ambiguous"]


def test_cli_cpu_form_with_physics_flag(rc, tmp_path):
    prefix = str(tmp_path / "out_")
    r = _dry(golden_path("sddm3.mtx"), "8", prefix, "1")
    assert r.returncode == 0, r.stderr
    cfg = json.loads(r.stdout.strip().splitlines()[-1])
    assert cfg["form"] == "cpu" and cfg["threads"] == 8 and cfg["dump"] == prefix
    assert cfg["kind"] == rc.INPUT_GROUNDED and cfg["check_solve"] == 0


def test_cli_cpu_form_without_flag_and_empty_prefix(rc):
    r = _dry(golden_path("grid3x3.mtx"), "4", "")
    assert r.returncode == 0, r.stderr
    cfg = json.loads(r.stdout.strip().splitlines()[-1])
    assert cfg["form"] == "cpu" and cfg["threads"] == 4 and cfg["kind"] == rc.INPUT_LAPLACIAN


def test_cli_ambiguous_tolerance_one(rc):
    # "<mtx> 8 1 1": GPU form with tol 1, or CPU form with prefix "1" + SDDM flag
    r = _dry(golden_path("grid3x3.mtx"), "8", "1", "1")
    assert r.returncode == 0, r.stderr
    cfg = json.loads(r.stdout.strip().splitlines()[-1])
    assert cfg["form"] == "gpu" and cfg["tol"] == 1.0 and cfg["ambiguous"] == 1
    assert "both forms" in r.stderr
    r = _dry(golden_path("sddm3.mtx"), "8", "1", "1", "--cpu-form")
    cfg = json.loads(r.stdout.strip().splitlines()[-1])
    assert cfg["form"] == "cpu" and cfg["dump"] == "1" and cfg["kind"] == rc.INPUT_GROUNDED


@pytest.mark.parametrize("args", [("x.mtx",), ("x.mtx", "a", "1", "1e-6"), ("x.mtx", "8", "p", "2"),
                                  ("x.mtx", "8", "1", "1e-6", "extra"), ("x.mtx", "8", "1", "1e-6", "--bogus")])
def test_cli_usage_errors(rc, args):
    r = _dry(*args)
    assert r.returncode == 2


def test_cli_orders_and_writes_the_prepared_matrix(rc, tmp_path):
    out = str(tmp_path / "ordered.mtx")
    r = _dry(golden_path("grid3x3.mtx"), "512", "0", "1e-6", "--order", "nd", "--write-ordered", out)
    assert r.returncode == 0, r.stderr
    assert json.loads(r.stdout.strip().splitlines()[-1])["order"] == rc.ORDER_ND
    assert open(out).readline().startswith("%%MatrixMarket")


@pytest.mark.parametrize("dims", [(48, 48), (14, 14, 14)])
def test_bisection_locality_order(rc, wl, dims):
    """The sweeps' within-level order: a permutation of the input vertices
    that depends on the graph only (not on the elimination ordering nor on
    the host thread count), with the input numbering when switched off."""
    n, r, c, v = wl.grid_laplacian(dims)
    s = _host(rc)
    s.set_triplets(n, r, c, v)
    s.order(rc.ORDER_NNZ_SORT, 1)
    rank0, _ = s.locality()
    np.testing.assert_array_equal(rank0, np.arange(n))              # off: input numbering
    s.set_locality(rc.LOCALITY_BISECTION)
    rank, secs = s.locality()
    assert sorted(rank.tolist()) == list(range(n)) and secs >= 0.0
    assert not np.array_equal(rank, np.arange(n))
    s.order(rc.ORDER_RANDOM, 7)                                         # another elimination ordering
    np.testing.assert_array_equal(s.locality()[0], rank)
    t = _host(rc, 1)
    t.set_locality(rc.LOCALITY_BISECTION)                               # set before the matrix: computed at order
    t.set_triplets(n, r, c, v)
    t.order(rc.ORDER_NNZ_SORT, 1)
    np.testing.assert_array_equal(t.locality()[0], rank)                # one host thread, same order
    s.set_locality(rc.LOCALITY_INPUT)
    np.testing.assert_array_equal(s.locality()[0], np.arange(n))


def test_bisection_locality_order_batch(rc, wl):
    """A batch: every system is bisected on its own; its ranks stay inside its
    row range and equal the ranks of the same system loaded alone."""
    systems = [wl.random_geometric_graph(k, 5 + k) for k in (300, 1000, 450)]
    s = _host(rc)
    s.set_batch(systems)
    s.set_locality(rc.LOCALITY_BISECTION)
    s.order(rc.ORDER_NNZ_SORT, 1)
    rank, _ = s.locality()
    off = 0
    for n, r, c, v in systems:
        part = rank[off:off + n]
        assert sorted(part.tolist()) == list(range(off, off + n))
        t = _host(rc)
        t.set_triplets(n, r, c, v)
        t.set_locality(rc.LOCALITY_BISECTION)
        t.order(rc.ORDER_NNZ_SORT, 1)
        np.testing.assert_array_equal(part - off, t.locality()[0])
        off += n
