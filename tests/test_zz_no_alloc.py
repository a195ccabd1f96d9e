"""rc_solve_noalloc against rc_solve through the C ABI: same factor, same
right-hand side, device-resident b / x (device buffers made by the test with
the HIP runtime), two calls in a row with the second system smaller."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    H2D, D2H = 1, 2

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p.value

    def upload(self, dev, a):
        assert self.lib.hipMemcpy(dev, a.ctypes.data, a.nbytes, self.H2D) == 0

    def download(self, a, dev):
        assert self.lib.hipMemcpy(a.ctypes.data, dev, a.nbytes, self.D2H) == 0

    def free(self, dev):
        assert self.lib.hipFree(dev) == 0


def _factored(rc, wl, dims, weights, prec):
    s = rc.Solver(device=0, seed=11, precision=prec)
    n, r, c, v = wl.grid_laplacian(dims, weights, seed=4)
    s.set_triplets(n, r, c, v)
    s.order(rc.ORDER_NNZ_SORT, 1)
    s.factor()
    return s, n


@pytest.mark.parametrize("prec", [64, 32])
def test_noalloc_equals_rc_solve_on_two_calls(rc, wl, prec):
    hip = _Hip()
    # first call on the larger system, second on a smaller one
    for dims in ((40, 40, 24), (24, 24)):
        s, n = _factored(rc, wl, dims, "loguniform", prec)
        print(f"noalloc: dims={dims} prec={prec} n={n}", flush=True)
        b = s.make_rhs(rc.RHS_NORMAL_PROJECTED, 42)
        x_ref, st_ref = s.solve(b, tol=1e-8)
        db, dx = hip.alloc(b.nbytes), hip.alloc(b.nbytes)
        try:
            hip.upload(db, b)
            hip.upload(dx, np.full(n, np.nan))
            st = s.solve_noalloc(db, dx, tol=1e-8)
            x = np.empty(n)
            hip.download(x, dx)
        finally:
            hip.free(db)
            hip.free(dx)
        assert st["iterations"] == st_ref["iterations"], (st, st_ref)
        assert st["converged"] == st_ref["converged"] == 1
        assert np.linalg.norm(x - x_ref) <= 1e-12 * np.linalg.norm(x_ref)
        assert abs(st["rel_res_true"] - st_ref["rel_res_true"]) <= 1e-6 * st_ref["rel_res_true"] + 1e-300
        s.close()


def test_noalloc_rejects_systems_beyond_the_workspace(rc, wl):
    s, n = _factored(rc, wl, (1100, 1000), "unit", 64)
    with pytest.raises(rc.RcholError) as e:
        s.solve_noalloc(0x1000, 0x1000)         # never dereferenced: the size check comes first
    assert e.value.status == rc.ERR_ARG
    assert "workspace" in s.lib.rc_last_error(s.h).decode()
    s.close()


def test_noalloc_rejects_batches_and_unfactored_handles(rc, wl):
    s = rc.Solver(device=0)
    s.set_triplets(*wl.grid_laplacian((6, 6)))
    with pytest.raises(rc.RcholError) as e:
        s.solve_noalloc(0x1000, 0x1000)
    assert e.value.status == rc.ERR_STATE
    s.set_batch([wl.grid_laplacian((5, 5)), wl.grid_laplacian((4, 6))])
    s.order(rc.ORDER_NNZ_SORT, 1)
    s.factor()
    with pytest.raises(rc.RcholError) as e:
        s.solve_noalloc(0x1000, 0x1000)
    assert e.value.status == rc.ERR_ARG
    s.close()
