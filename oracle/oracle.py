"""ctypes wrapper of liboracle.so, the CPU oracle (test infrastructure only).

Only the tests, the smoke check and the CPU-baseline leg of the benchmark use
this module, and only as the checker.  See rchol_oracle.h for the reference
lines each function restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")

MUT_NONE, MUT_FILL_WEIGHT, MUT_SAMPLE_SHIFT, MUT_NO_MERGE, MUT_PIVOT, MUT_PRECOND_SIGN = 0, 1, 2, 3, 4, 5
MUT_IC0_NO_PIVOT = 6
STOP_REFERENCE, STOP_TRUE_RES = 0, 1

_i64 = np.ctypeslib.ndpointer(np.int64, flags="C")
_i32 = np.ctypeslib.ndpointer(np.int32, flags="C")
_f64 = np.ctypeslib.ndpointer(np.float64, flags="C")
_u32 = np.ctypeslib.ndpointer(np.uint32, flags="C")


class FactorStats(ctypes.Structure):
    _fields_ = [("nnz_off", ctypes.c_int64), ("fills", ctypes.c_int64), ("gathered", ctypes.c_int64),
                ("max_degree", ctypes.c_int64), ("nlevels", ctypes.c_int32), ("zero_pivots", ctypes.c_int32)]


class PcgStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("converged", ctypes.c_int32), ("rz0", ctypes.c_double),
                ("rz", ctypes.c_double), ("rel_res_true", ctypes.c_double), ("rel_res_rec", ctypes.c_double)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    L = ctypes.CDLL(LIB)
    L.orc_philox4x32.argtypes = [_u32, _u32, _u32]
    L.orc_philox4x32.restype = None
    L.orc_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
    L.orc_uniform.restype = ctypes.c_double
    L.orc_normal.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32]
    L.orc_normal.restype = ctypes.c_double
    L.orc_laplacian_from_triplets.argtypes = [ctypes.c_int, ctypes.c_int64, _i32, _i32, _f64, _i64, _i32, _f64]
    L.orc_laplacian_from_triplets.restype = ctypes.c_int64
    L.orc_dependency_counts.argtypes = [ctypes.c_int, _i64, _i32, _i32]
    L.orc_dependency_counts.restype = None
    L.orc_ground_append.argtypes = [ctypes.c_int, _i64, _i32, _f64, _i64, _i32, _f64]
    L.orc_ground_append.restype = ctypes.c_int64
    L.orc_trim_last.argtypes = [ctypes.c_int, _i64, _i32, _f64, _i64, _i32, _f64]
    L.orc_trim_last.restype = ctypes.c_int64
    L.orc_random_perm.argtypes = [ctypes.c_int, ctypes.c_uint64, _i32]
    L.orc_random_perm.restype = None
    L.orc_nnz_sort_perm.argtypes = [ctypes.c_int, _i64, ctypes.c_uint64, _i32]
    L.orc_nnz_sort_perm.restype = None
    L.orc_permute.argtypes = [ctypes.c_int, _i64, _i32, _f64, _i32, _i64, _i32, _f64]
    L.orc_permute.restype = ctypes.c_int64
    L.orc_factor.argtypes = [ctypes.c_int, _i64, _i32, _f64, ctypes.c_uint64, ctypes.c_int, ctypes.c_int64, _i64,
                             _i32, _f64, _f64, _i32, ctypes.POINTER(FactorStats)]
    L.orc_factor.restype = ctypes.c_int
    L.orc_factor_trim_last.argtypes = [ctypes.c_int, _i64, _i32, _f64, _i64, _i32, _f64]
    L.orc_factor_trim_last.restype = ctypes.c_int64
    L.orc_spmv.argtypes = [ctypes.c_int, _i64, _i32, _f64, _f64, _f64]
    L.orc_spmv.restype = None
    L.orc_precond.argtypes = [ctypes.c_int, _i64, _i32, _f64, _f64, ctypes.c_int, ctypes.c_int, _f64, _f64]
    L.orc_precond.restype = None
    L.orc_pcg.argtypes = [ctypes.c_int, _i64, _i32, _f64, _i64, _i32, _f64, _f64, ctypes.c_int, _f64, _f64,
                          ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.POINTER(PcgStats)]
    L.orc_pcg.restype = ctypes.c_int
    L.orc_ic0.argtypes = [ctypes.c_int, _i64, _i32, _f64, ctypes.c_int, ctypes.c_int64, _i64, _i32, _f64, _f64, _i32]
    L.orc_ic0.restype = ctypes.c_int
    L.orc_normal_vector.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_uint32, _f64]
    L.orc_normal_vector.restype = None
    L.orc_factor_omp.argtypes = [ctypes.c_int, _i64, _i32, _f64, ctypes.c_uint64, ctypes.c_int, ctypes.c_int64,
                                 _i64, _i32, _f64, _f64, _i32, ctypes.POINTER(FactorStats)]
    L.orc_factor_omp.restype = ctypes.c_int
    L.orc_pcg_omp.argtypes = [ctypes.c_int, _i64, _i32, _f64, _i64, _i32, _f64, _f64, _i32, _f64, _f64,
                              ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(PcgStats)]
    L.orc_pcg_omp.restype = ctypes.c_int
    _lib = L
    return L


class Csc:
    """Square CSC matrix (int64 pointers, int32 rows)."""

    def __init__(self, n, ptr, ind, val):
        self.n = int(n)
        self.ptr = np.ascontiguousarray(ptr, dtype=np.int64)
        self.ind = np.ascontiguousarray(ind, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)

    @property
    def nnz(self):
        return int(self.ptr[-1])

    def dense(self):
        a = np.zeros((self.n, self.n))
        for j in range(self.n):
            for p in range(self.ptr[j], self.ptr[j + 1]):
                a[self.ind[p], j] += self.val[p]
        return a

    def scipy(self):
        import scipy.sparse as sp
        return sp.csc_matrix((self.val, self.ind, self.ptr), shape=(self.n, self.n))


def philox4x32(ctr, key):
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32(np.asarray(ctr, dtype=np.uint32), np.asarray(key, dtype=np.uint32), out)
    return out


def uniform(seed, a, b, tag=0):
    return lib().orc_uniform(seed, a, b, tag)


def normal(seed, index, stream=0):
    return lib().orc_normal(seed, index, stream)


def laplacian(n, rows, cols, vals):
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    m = len(rows)
    ptr = np.zeros(n + 1, np.int64)
    ind = np.zeros(m + n, np.int32)
    val = np.zeros(m + n, np.float64)
    nnz = lib().orc_laplacian_from_triplets(n, m, rows, cols, vals, ptr, ind, val)
    if nnz < 0:
        raise ValueError("index out of range")
    return Csc(n, ptr, ind[:nnz], val[:nnz])


def dependency_counts(a: Csc):
    dep = np.zeros(a.n, np.int32)
    lib().orc_dependency_counts(a.n, a.ptr, a.ind, dep)
    return dep


def ground_append(a: Csc):
    """Returns (Csc, appended) or raises ValueError if not dominant."""
    cap = a.nnz + 2 * a.n + 1
    ptr = np.zeros(a.n + 2, np.int64)
    ind = np.zeros(cap, np.int32)
    val = np.zeros(cap, np.float64)
    q = lib().orc_ground_append(a.n, a.ptr, a.ind, a.val, ptr, ind, val)
    if q == -2:
        raise ValueError("not diagonally dominant")
    if q == 0:
        return a, False
    return Csc(a.n + 1, ptr, ind[:q], val[:q]), True


def trim_last(a: Csc):
    ptr = np.zeros(a.n, np.int64)
    ind = np.zeros(max(a.nnz, 1), np.int32)
    val = np.zeros(max(a.nnz, 1), np.float64)
    q = lib().orc_trim_last(a.n, a.ptr, a.ind, a.val, ptr, ind, val)
    return Csc(a.n - 1, ptr, ind[:q], val[:q])


def random_perm(n, seed):
    p = np.zeros(n, np.int32)
    lib().orc_random_perm(n, seed, p)
    return p


def nnz_sort_perm(a: Csc, seed):
    p = np.zeros(a.n, np.int32)
    lib().orc_nnz_sort_perm(a.n, a.ptr, seed, p)
    return p


def permute(a: Csc, perm):
    perm = np.ascontiguousarray(perm, dtype=np.int32)
    ptr = np.zeros(a.n + 1, np.int64)
    ind = np.zeros(max(a.nnz, 1), np.int32)
    val = np.zeros(max(a.nnz, 1), np.float64)
    q = lib().orc_permute(a.n, a.ptr, a.ind, a.val, perm, ptr, ind, val)
    return Csc(a.n, ptr, ind[:q], val[:q])


class Factor:
    """Off-diagonal L^T in CSR (row k = column k of L, values -w/S), pivots, levels."""

    def __init__(self, n, ptr, ind, val, diag, level, stats):
        self.n, self.ptr, self.ind, self.val = n, ptr, ind, val
        self.diag, self.level, self.stats = diag, level, stats

    def dense_L(self):
        L = np.eye(self.n)
        for k in range(self.n):
            for p in range(self.ptr[k], self.ptr[k + 1]):
                L[self.ind[p], k] = self.val[p]
        return L


def factor(a: Csc, seed=1, mutation=MUT_NONE, cap=None):
    n = a.n
    if cap is None:
        cap = max(16 * a.nnz, 1024)
    ptr = np.zeros(n + 1, np.int64)
    ind = np.zeros(cap, np.int32)
    val = np.zeros(cap, np.float64)
    diag = np.zeros(n, np.float64)
    level = np.zeros(n, np.int32)
    st = FactorStats()
    rc = lib().orc_factor(n, a.ptr, a.ind, a.val, seed, mutation, cap, ptr, ind, val, diag, level, ctypes.byref(st))
    if rc != 0:
        raise RuntimeError("oracle factor capacity exceeded")
    q = int(ptr[-1])
    stats = {k: getattr(st, k) for k, _ in st._fields_}
    return Factor(n, ptr, ind[:q].copy(), val[:q].copy(), diag, level, stats)


def normal_vector(seed, n, stream=0):
    """[normal(seed, i, stream) for i < n], computed in C."""
    out = np.empty(int(n), np.float64)
    lib().orc_normal_vector(seed, int(n), stream, out)
    return out


def factor_omp(a: Csc, seed=1, threads=0, cap=None):
    """orc_factor scheduled on `threads` OpenMP threads (0 = all); bitwise equal to factor()."""
    n = a.n
    if cap is None:
        cap = max(16 * a.nnz, 1024)
    ptr = np.zeros(n + 1, np.int64)
    ind = np.zeros(cap, np.int32)
    val = np.zeros(cap, np.float64)
    diag = np.zeros(n, np.float64)
    level = np.zeros(n, np.int32)
    st = FactorStats()
    rc = lib().orc_factor_omp(n, a.ptr, a.ind, a.val, seed, int(threads), cap, ptr, ind, val, diag, level,
                              ctypes.byref(st))
    if rc != 0:
        raise RuntimeError("oracle factor capacity exceeded")
    q = int(ptr[-1])
    stats = {k: getattr(st, k) for k, _ in st._fields_}
    return Factor(n, ptr, ind[:q], val[:q], diag, level, stats)


def pcg_omp(a: Csc, f: Factor, b, tol=1e-6, maxit=300, threads=0):
    """The reference PCG (FP64, reference stop rule) with OpenMP kernels and
    level-scheduled triangular solves."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(a.n)
    st = PcgStats()
    setup = ctypes.c_double()
    level = np.ascontiguousarray(f.level[:a.n], dtype=np.int32)
    rc = lib().orc_pcg_omp(a.n, a.ptr, a.ind, a.val, f.ptr, f.ind, f.val, f.diag, level, b, x, tol, maxit,
                           int(threads), ctypes.byref(setup), ctypes.byref(st))
    out = {k: getattr(st, k) for k, _ in st._fields_}
    out["status"] = rc
    out["setup_s"] = setup.value
    return x, out


def ic0(a: Csc, mutation=MUT_NONE):
    """IC(0) L D L^T of a (pattern of its lower triangle), orc_factor's layout."""
    n = a.n
    cap = max(a.nnz, 1)
    ptr = np.zeros(n + 1, np.int64)
    ind = np.zeros(cap, np.int32)
    val = np.zeros(cap, np.float64)
    diag = np.zeros(n, np.float64)
    level = np.zeros(n, np.int32)
    if lib().orc_ic0(n, a.ptr, a.ind, a.val, mutation, cap, ptr, ind, val, diag, level) != 0:
        raise RuntimeError("ic0 capacity exceeded")
    q = int(ptr[-1])
    return Factor(n, ptr, ind[:q].copy(), val[:q].copy(), diag, level, {})


def factor_trim_last(f: Factor):
    n = f.n
    ptr = np.zeros(n, np.int64)
    ind = np.zeros(max(len(f.ind), 1), np.int32)
    val = np.zeros(max(len(f.ind), 1), np.float64)
    q = lib().orc_factor_trim_last(n, f.ptr, f.ind, f.val, ptr, ind, val)
    return Factor(n - 1, ptr, ind[:q].copy(), val[:q].copy(), f.diag[:n - 1].copy(), f.level[:n - 1].copy(), f.stats)


def spmv(a: Csc, x):
    y = np.zeros(a.n)
    lib().orc_spmv(a.n, a.ptr, a.ind, a.val, np.ascontiguousarray(x, dtype=np.float64), y)
    return y


def precond(f: Factor, r, precision=64, mutation=MUT_NONE):
    z = np.zeros(f.n)
    lib().orc_precond(f.n, f.ptr, f.ind, f.val, f.diag, precision, mutation,
                      np.ascontiguousarray(r, dtype=np.float64), z)
    return z


def pcg(a: Csc, f: Factor, b, tol=1e-6, maxit=300, precision=64, stop_rule=STOP_REFERENCE):
    x = np.zeros(a.n)
    st = PcgStats()
    rc = lib().orc_pcg(a.n, a.ptr, a.ind, a.val, f.ptr, f.ind, f.val, f.diag, precision,
                       np.ascontiguousarray(b, dtype=np.float64), x, tol, maxit, stop_rule, ctypes.byref(st))
    stats = {k: getattr(st, k) for k, _ in st._fields_}
    stats["rc"] = rc
    return x, stats


def prepare_system(n, rows, cols, vals, kind="laplacian", order="nnz-sort", order_seed=1):
    """The whole input path on the oracle: Laplacian build, ordering, ground
    append (SDDM) -> (factored Laplacian, PCG system, perm, trim flag)."""
    a = laplacian(n, rows, cols, vals)
    if kind == "sddm":
        perm = nnz_sort_perm(a, order_seed) if order == "nnz-sort" else (
            random_perm(n, order_seed) if order == "random" else np.arange(n, dtype=np.int32))
        sp = permute(a, perm)
        fac, appended = ground_append(sp)
        return fac, sp, perm, appended
    perm = nnz_sort_perm(a, order_seed) if order == "nnz-sort" else (
        random_perm(n, order_seed) if order == "random" else np.arange(n, dtype=np.int32))
    lap = permute(a, perm)
    return lap, lap, perm, False
