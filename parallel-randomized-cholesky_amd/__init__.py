"""Python binding of librchol.so, the MI355X randomized-Cholesky + PCG library.

A thin ctypes layer over the C ABI declared in include/rchol.h, used by the
tests, the benchmark and the smoke check.  The compute path is the HIP library;
there is no Python or CPU fallback: if librchol.so is missing or no HIP device
is present, the constructor raises.

Mirrors the reference driver surface (gpu_implementation/driver.cu:2148-2157,
driver_physics.cu:2120-2129, cpu_implementation/driver_local.cpp:1325-1347):
Matrix Market input, SDDM->Laplacian ground append, orderings, thread-block
count, tolerance, factor dump.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librchol.so")

# rc_status
OK = 0
ERR_ARG, ERR_IO, ERR_FORMAT, ERR_NOT_DOMINANT, ERR_STATE, ERR_HIP, ERR_POOL = -1, -2, -3, -4, -5, -6, -7
ERR_TIMEOUT, ERR_BREAKDOWN, ERR_NO_DEVICE, ERR_NOMEM, ERR_NOT_SYMMETRIC, ERR_MAXIT = -8, -9, -10, -11, -12, -13

INPUT_LAPLACIAN, INPUT_GROUNDED, INPUT_SDDM = 0, 1, 2
ORDER_AS_IS, ORDER_NNZ_SORT, ORDER_RANDOM, ORDER_AMD, ORDER_LEARNED, ORDER_ND = 0, 1, 2, 3, 4, 5
LOCALITY_INPUT, LOCALITY_BISECTION = 0, 1
STOP_REFERENCE, STOP_TRUE_RESIDUAL = 0, 1
RHS_ZERO_SUM_UNIFORM, RHS_NORMAL, RHS_NORMAL_PROJECTED = 0, 1, 2
KERNEL_SPMV, KERNEL_TRSV_FWD, KERNEL_TRSV_BWD, KERNEL_PCG_ITER, KERNEL_COPY = 0, 1, 2, 3, 4
KERNEL_MULTI_SPMM, KERNEL_MULTI_TRSV_FWD, KERNEL_MULTI_TRSV_BWD = 5, 6, 7

# Every symbol include/rchol.h declares.
EXPORTS = (
    "rc_default_options", "rc_create", "rc_destroy", "rc_last_error", "rc_status_string",
    "rc_load_mtx", "rc_set_triplets", "rc_get_size", "rc_order", "rc_get_permutation",
    "rc_set_gnn_weights", "rc_factor", "rc_get_factor", "rc_dump_factor", "rc_load_factor",
    "rc_solve", "rc_make_rhs", "rc_set_reference_solution", "rc_set_rhs", "rc_solve_resident",
    "rc_get_solution", "rc_time_kernel", "rc_get_prepared_csc", "rc_write_prepared_mtx", "rc_solve_multi",
    "rc_set_permutation", "rc_gnn_weight_count", "rc_gnn_scores", "rc_gnn_gradient",
    "rc_set_batch", "rc_get_batch", "rc_solve_batch", "rc_get_batch_stats", "rc_autotune",
    "rc_set_rhs_multi", "rc_solve_multi_resident", "rc_get_solution_multi", "rc_solve_noalloc",
    "rc_get_multi_stats", "rc_solve_group", "rc_set_solve_options", "rc_factor_ic0",
    "rc_set_locality", "rc_get_locality",
)


class TuneStats(ctypes.Structure):
    _fields_ = [("tail_max_rows", ctypes.c_int32), ("sweep_blocks", ctypes.c_int32),
                ("configs_tried", ctypes.c_int32), ("sweeps_ms_before", ctypes.c_double),
                ("sweeps_ms_after", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class RcholError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"rchol error {status}: {msg}")
        self.status = status


class Options(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int),
        ("seed", ctypes.c_uint64),
        ("num_blocks", ctypes.c_int),
        ("factor_precision", ctypes.c_int),
        ("stop_rule", ctypes.c_int),
        ("maxit", ctypes.c_int),
        ("pool_factor", ctypes.c_double),
        ("verbose", ctypes.c_int),
        ("host_threads", ctypes.c_int),
    ]


class FactorStats(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64), ("nnz_lower", ctypes.c_int64), ("nnz_factor", ctypes.c_int64),
        ("fills", ctypes.c_int64), ("max_degree", ctypes.c_int64), ("nlevels", ctypes.c_int32),
        ("zero_pivots", ctypes.c_int32), ("factor_ms", ctypes.c_double), ("assemble_ms", ctypes.c_double),
        ("bytes_factor", ctypes.c_double), ("pool_regrowths", ctypes.c_int32), ("prepare_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SolveStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32), ("converged", ctypes.c_int32), ("rel_res_true", ctypes.c_double),
        ("rel_res_precond", ctypes.c_double), ("pcg_ms", ctypes.c_double), ("setup_ms", ctypes.c_double),
        ("err_vs_oracle", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load librchol.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RcholError(ERR_NO_DEVICE, f"{path} not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    lib.rc_default_options.argtypes = [P(Options)]
    lib.rc_default_options.restype = None
    lib.rc_create.argtypes = [P(vp), P(Options)]
    lib.rc_destroy.argtypes = [vp]
    lib.rc_last_error.argtypes = [vp]
    lib.rc_last_error.restype = ctypes.c_char_p
    lib.rc_status_string.argtypes = [ctypes.c_int]
    lib.rc_status_string.restype = ctypes.c_char_p
    lib.rc_load_mtx.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
    lib.rc_set_triplets.argtypes = [vp, ctypes.c_int, ctypes.c_int64, vp, vp, vp, ctypes.c_int]
    lib.rc_get_size.argtypes = [vp, P(ctypes.c_int64), P(ctypes.c_int64)]
    lib.rc_order.argtypes = [vp, ctypes.c_int, ctypes.c_uint64]
    lib.rc_get_permutation.argtypes = [vp, vp]
    lib.rc_set_gnn_weights.argtypes = [vp, vp, ctypes.c_int64]
    lib.rc_factor.argtypes = [vp, P(FactorStats)]
    lib.rc_get_factor.argtypes = [vp, P(ctypes.c_int64), P(ctypes.c_int64), vp, vp, vp, vp]
    lib.rc_dump_factor.argtypes = [vp, ctypes.c_char_p]
    lib.rc_load_factor.argtypes = [vp, ctypes.c_char_p]
    lib.rc_solve.argtypes = [vp, vp, vp, ctypes.c_double, P(SolveStats)]
    lib.rc_make_rhs.argtypes = [vp, ctypes.c_int, ctypes.c_uint64, vp]
    lib.rc_set_reference_solution.argtypes = [vp, vp]
    lib.rc_set_rhs.argtypes = [vp, vp]
    lib.rc_solve_resident.argtypes = [vp, ctypes.c_double, P(SolveStats)]
    lib.rc_get_solution.argtypes = [vp, vp]
    lib.rc_time_kernel.argtypes = [vp, ctypes.c_int, ctypes.c_int, P(ctypes.c_double), P(ctypes.c_double)]
    lib.rc_get_prepared_csc.argtypes = [vp, P(ctypes.c_int64), P(ctypes.c_int64), vp, vp, vp]
    lib.rc_write_prepared_mtx.argtypes = [vp, ctypes.c_char_p]
    lib.rc_solve_multi.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_double, P(SolveStats)]
    lib.rc_set_permutation.argtypes = [vp, vp]
    lib.rc_set_locality.argtypes = [vp, ctypes.c_int]
    lib.rc_get_locality.argtypes = [vp, vp, P(ctypes.c_double)]
    lib.rc_gnn_weight_count.argtypes = []
    lib.rc_gnn_weight_count.restype = ctypes.c_int64
    lib.rc_gnn_scores.argtypes = [vp, vp, vp]
    lib.rc_gnn_gradient.argtypes = [vp, vp, vp, vp]
    lib.rc_set_batch.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp, vp]
    lib.rc_get_batch.argtypes = [vp, P(ctypes.c_int), vp]
    lib.rc_solve_batch.argtypes = [vp, vp, vp, ctypes.c_double, vp]
    lib.rc_get_batch_stats.argtypes = [vp, vp]
    lib.rc_autotune.argtypes = [vp, ctypes.c_int, P(TuneStats)]
    lib.rc_set_rhs_multi.argtypes = [vp, ctypes.c_int, vp]
    lib.rc_solve_multi_resident.argtypes = [vp, ctypes.c_double, vp]
    lib.rc_get_solution_multi.argtypes = [vp, vp]
    lib.rc_solve_noalloc.argtypes = [vp, vp, vp, ctypes.c_double, P(SolveStats)]
    lib.rc_get_multi_stats.argtypes = [vp, vp]
    lib.rc_solve_group.argtypes = [ctypes.c_int, vp, ctypes.c_double, vp]
    lib.rc_set_solve_options.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    lib.rc_factor_ic0.argtypes = [vp, P(FactorStats)]
    _lib = lib
    return lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class Solver:
    """One rc_handle: one HIP stream and the device memory of one system.
    device=-1 makes a host-only handle (input, ordering, RHS; no factor/solve)."""

    def __init__(self, device=0, seed=1, num_blocks=0, precision=64, stop_rule=STOP_REFERENCE, maxit=300,
                 pool_factor=4.0, verbose=0, host_threads=0):
        self.lib = load_library()
        o = Options()
        self.lib.rc_default_options(ctypes.byref(o))
        o.device, o.seed, o.num_blocks = device, seed, num_blocks
        o.factor_precision, o.stop_rule, o.maxit = precision, stop_rule, maxit
        o.pool_factor, o.verbose, o.host_threads = pool_factor, verbose, host_threads
        h = ctypes.c_void_p()
        st = self.lib.rc_create(ctypes.byref(h), ctypes.byref(o))
        if st != OK:
            raise RcholError(st, self.lib.rc_status_string(st).decode())
        self.h = h
        self.precision = precision

    # -- plumbing
    def _check(self, st):
        if st != OK:
            raise RcholError(st, self.lib.rc_last_error(self.h).decode(errors="replace"))

    def close(self):
        if getattr(self, "h", None):
            self.lib.rc_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- input
    def load_mtx(self, path, kind=INPUT_LAPLACIAN):
        self._check(self.lib.rc_load_mtx(self.h, str(path).encode(), kind))

    def set_triplets(self, n, rows, cols, vals, kind=INPUT_LAPLACIAN):
        r = np.ascontiguousarray(rows, dtype=np.int32)
        c = np.ascontiguousarray(cols, dtype=np.int32)
        v = np.ascontiguousarray(vals, dtype=np.float64)
        self._check(self.lib.rc_set_triplets(self.h, int(n), len(r), _ptr(r), _ptr(c), _ptr(v), kind))

    def set_batch(self, systems):
        """systems: list of (n, rows, cols, vals) with system-local indices."""
        ns = np.array([int(t[0]) for t in systems], dtype=np.int32)
        ms = np.array([len(t[1]) for t in systems], dtype=np.int64)
        r = np.ascontiguousarray(np.concatenate([np.asarray(t[1]) for t in systems]), dtype=np.int32)
        c = np.ascontiguousarray(np.concatenate([np.asarray(t[2]) for t in systems]), dtype=np.int32)
        v = np.ascontiguousarray(np.concatenate([np.asarray(t[3]) for t in systems]), dtype=np.float64)
        self._check(self.lib.rc_set_batch(self.h, len(systems), _ptr(ns), _ptr(ms), _ptr(r), _ptr(c), _ptr(v)))

    def batch(self):
        """(number of systems, offsets[nsys + 1]) of the loaded input."""
        k = ctypes.c_int()
        self._check(self.lib.rc_get_batch(self.h, ctypes.byref(k), None))
        off = np.zeros(k.value + 1, dtype=np.int64)
        self._check(self.lib.rc_get_batch(self.h, ctypes.byref(k), _ptr(off)))
        return k.value, off

    def solve_batch(self, b, tol=1e-6, allow_breakdown=False):
        """One PCG per system; returns x and the list of per-system stats."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(b)
        k, _ = self.batch()
        st = (SolveStats * k)()
        rc = self.lib.rc_solve_batch(self.h, _ptr(b), _ptr(x), float(tol), st)
        if not (rc == OK or (allow_breakdown and rc == ERR_BREAKDOWN)):
            self._check(rc)
        return x, [st[q].as_dict() for q in range(k)]

    def batch_stats(self):
        k, _ = self.batch()
        st = (SolveStats * k)()
        self._check(self.lib.rc_get_batch_stats(self.h, st))
        return [st[q].as_dict() for q in range(k)]

    def size(self):
        n, nnz = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.rc_get_size(self.h, ctypes.byref(n), ctypes.byref(nnz)))
        return n.value, nnz.value

    def prepared_csc(self):
        """The ordered (ground-appended for SDDM input) Laplacian that is factored."""
        n, nnz = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.rc_get_prepared_csc(self.h, ctypes.byref(n), ctypes.byref(nnz), None, None, None))
        ptr = np.empty(n.value + 1, dtype=np.int64)
        ind = np.empty(nnz.value, dtype=np.int32)
        val = np.empty(nnz.value, dtype=np.float64)
        self._check(self.lib.rc_get_prepared_csc(self.h, ctypes.byref(n), ctypes.byref(nnz), _ptr(ptr), _ptr(ind),
                                                 _ptr(val)))
        return n.value, ptr, ind, val

    def write_prepared_mtx(self, path):
        self._check(self.lib.rc_write_prepared_mtx(self.h, str(path).encode()))

    # -- ordering
    def order(self, method=ORDER_NNZ_SORT, seed=1):
        self._check(self.lib.rc_order(self.h, method, seed))

    def permutation(self):
        n, _ = self.size()
        p = np.empty(n, dtype=np.int32)
        self._check(self.lib.rc_get_permutation(self.h, _ptr(p)))
        return p

    def set_permutation(self, perm):
        p = np.ascontiguousarray(perm, dtype=np.int32)
        self._check(self.lib.rc_set_permutation(self.h, _ptr(p)))

    def set_locality(self, mode=LOCALITY_BISECTION):
        """Within-level row order of the sweeps: input numbering or recursive
        bisection of the graph (computed once per matrix on the host)."""
        self._check(self.lib.rc_set_locality(self.h, mode))

    def locality(self):
        """(rank per input vertex, host seconds of the last bisection)."""
        n, _ = self.size()
        r = np.empty(n, dtype=np.int32)
        sec = ctypes.c_double(0.0)
        self._check(self.lib.rc_get_locality(self.h, _ptr(r), ctypes.byref(sec)))
        return r, sec.value

    def gnn_scores(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        n, _ = self.size()
        s = np.empty(n, dtype=np.float32)
        self._check(self.lib.rc_gnn_scores(self.h, _ptr(w), _ptr(s)))
        return s

    def gnn_gradient(self, w, ds):
        w = np.ascontiguousarray(w, dtype=np.float32)
        ds = np.ascontiguousarray(ds, dtype=np.float32)
        dw = np.empty_like(w)
        self._check(self.lib.rc_gnn_gradient(self.h, _ptr(w), _ptr(ds), _ptr(dw)))
        return dw

    def set_gnn_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._check(self.lib.rc_set_gnn_weights(self.h, _ptr(w), len(w)))

    # -- factor
    def factor(self):
        st = FactorStats()
        self._check(self.lib.rc_factor(self.h, ctypes.byref(st)))
        return st.as_dict()

    def factor_ic0(self):
        """Incomplete Cholesky IC(0) of the prepared matrix as the preconditioner."""
        st = FactorStats()
        self._check(self.lib.rc_factor_ic0(self.h, ctypes.byref(st)))
        return st.as_dict()

    def get_factor(self):
        """CSR of L^T with unit diagonal first (reference layout) and the pivots."""
        n, nnz = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.rc_get_factor(self.h, ctypes.byref(n), ctypes.byref(nnz), None, None, None, None))
        rp = np.empty(n.value + 1, dtype=np.int64)
        ci = np.empty(nnz.value, dtype=np.int32)
        va = np.empty(nnz.value, dtype=np.float64)
        dg = np.empty(n.value, dtype=np.float64)
        self._check(self.lib.rc_get_factor(self.h, ctypes.byref(n), ctypes.byref(nnz), _ptr(rp), _ptr(ci), _ptr(va),
                                           _ptr(dg)))
        return rp, ci, va, dg

    def dump_factor(self, prefix):
        self._check(self.lib.rc_dump_factor(self.h, str(prefix).encode()))

    def load_factor(self, prefix):
        self._check(self.lib.rc_load_factor(self.h, str(prefix).encode()))

    # -- solve
    def make_rhs(self, kind=RHS_NORMAL_PROJECTED, seed=42):
        n, _ = self.size()
        b = np.empty(n, dtype=np.float64)
        self._check(self.lib.rc_make_rhs(self.h, kind, seed, _ptr(b)))
        return b

    def set_reference_solution(self, x_ref):
        if x_ref is None:
            self._check(self.lib.rc_set_reference_solution(self.h, None))
        else:
            self._xref = np.ascontiguousarray(x_ref, dtype=np.float64)
            self._check(self.lib.rc_set_reference_solution(self.h, _ptr(self._xref)))

    def solve(self, b, tol=1e-6, allow_breakdown=False):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(b)
        st = SolveStats()
        rc = self.lib.rc_solve(self.h, _ptr(b), _ptr(x), float(tol), ctypes.byref(st))
        if not (rc == OK or (allow_breakdown and rc == ERR_BREAKDOWN)):
            self._check(rc)
        return x, st.as_dict()

    def solve_multi(self, B, tol=1e-6, allow_breakdown=False):
        """k <= 64 right-hand sides (columns of B, shape (n, k)); returns X, stats list."""
        B = np.asarray(B, dtype=np.float64)
        n, k = B.shape
        Bc = np.ascontiguousarray(B.T)          # column-major for the C ABI
        Xc = np.empty_like(Bc)
        st = (SolveStats * k)()
        rc = self.lib.rc_solve_multi(self.h, k, _ptr(Bc), _ptr(Xc), float(tol), st)
        if not (rc == OK or (allow_breakdown and rc == ERR_BREAKDOWN)):
            self._check(rc)
        return Xc.T.copy(), [st[q].as_dict() for q in range(k)]

    def set_rhs_multi(self, B):
        """Upload k right-hand sides (columns of B, shape (n, k)) once."""
        B = np.asarray(B, dtype=np.float64)
        self._Bc = np.ascontiguousarray(B.T)          # column-major for the C ABI
        self._k = B.shape[1]
        self._check(self.lib.rc_set_rhs_multi(self.h, self._k, _ptr(self._Bc)))

    def solve_multi_resident(self, tol=1e-6, allow_breakdown=False):
        st = (SolveStats * self._k)()
        rc = self.lib.rc_solve_multi_resident(self.h, float(tol), st)
        if not (rc == OK or (allow_breakdown and rc == ERR_BREAKDOWN)):
            self._check(rc)
        return [st[q].as_dict() for q in range(self._k)]

    def multi_stats(self):
        st = (SolveStats * self._k)()
        self._check(self.lib.rc_get_multi_stats(self.h, st))
        return [st[q].as_dict() for q in range(self._k)]

    def get_solution_multi(self):
        n, _ = self.size()
        Xc = np.empty((self._k, n), dtype=np.float64)
        self._check(self.lib.rc_get_solution_multi(self.h, _ptr(Xc)))
        return Xc.T.copy()

    def solve_noalloc(self, b_dev, x_dev, tol=1e-6):
        """PCG on device-resident b / x (device addresses as ints); allocates nothing."""
        st = SolveStats()
        self._check(self.lib.rc_solve_noalloc(self.h, ctypes.c_void_p(b_dev), ctypes.c_void_p(x_dev), float(tol),
                                              ctypes.byref(st)))
        return st.as_dict()

    def set_solve_options(self, stop_rule=STOP_REFERENCE, maxit=0):
        """Stop rule and (maxit > 0) iteration limit of later solves."""
        self._check(self.lib.rc_set_solve_options(self.h, stop_rule, maxit))

    def set_rhs(self, b):
        self._b = np.ascontiguousarray(b, dtype=np.float64)
        self._check(self.lib.rc_set_rhs(self.h, _ptr(self._b)))

    def solve_resident(self, tol=1e-6):
        st = SolveStats()
        self._check(self.lib.rc_solve_resident(self.h, float(tol), ctypes.byref(st)))
        return st.as_dict()

    def get_solution(self):
        n, _ = self.size()
        x = np.empty(n, dtype=np.float64)
        self._check(self.lib.rc_get_solution(self.h, _ptr(x)))
        return x

    def autotune(self, reps=5):
        """Pick the sweep launch configuration for the factored structure."""
        st = TuneStats()
        self._check(self.lib.rc_autotune(self.h, int(reps), ctypes.byref(st)))
        return st.as_dict()

    def time_kernel(self, kernel, reps=10):
        ms, by = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.rc_time_kernel(self.h, kernel, reps, ctypes.byref(ms), ctypes.byref(by)))
        return ms.value, by.value


def solve_group(solvers, tol=1e-6, allow_breakdown=False):
    """rc_solve_group over Solver objects whose right-hand sides are resident
    (set_rhs or set_rhs_multi); returns one summary stats dict per solver."""
    lib = load_library()
    hs = (ctypes.c_void_p * len(solvers))(*[s.h.value for s in solvers])
    st = (SolveStats * len(solvers))()
    rc = lib.rc_solve_group(len(solvers), hs, float(tol), st)
    if not (rc == OK or (allow_breakdown and rc == ERR_BREAKDOWN)):
        raise RcholError(rc, lib.rc_last_error(solvers[0].h).decode(errors="replace"))
    return [st[i].as_dict() for i in range(len(solvers))]
