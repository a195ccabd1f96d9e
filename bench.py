#!/usr/bin/env python3
"""Benchmark: randomized-Cholesky factor + PCG to relative residual 1e-6 on one
MI355X, with the CPU restatement of the same algorithm timed beside it.

One step = rc_factor (elimination + assembly of the resident system) +
rc_solve_resident (PCG to tol with the right-hand side already in HBM).  Both
calls return after synchronising the library's stream; the timed loop is
additionally bracketed by hipDeviceSynchronize.

Default workload (the headline comparison): 3D 7-point grid Laplacian 256^3
with edge weights 10^U(-3,3) (seed 7), one N(0,1) right-hand side (seed 42)
projected orthogonal to the ones vector, tol 1e-6, reference stop rule.
Two configurations are timed in the same invocation:
  * learned ordering + FP32 factor and triangular solves, FP64 SpMV / dots /
    recurrence (the target; `value` is its time per step);
  * plain nnz-sorted ordering, FP64 everywhere (the comparison).
Both are also run with the true-residual stop rule (||b - A x|| / ||b|| <=
1e-6) so the speed-up is stated at equal accuracy too, and the FP32-factor
solution's error against an FP64 reference solve is reported.

Other workloads (--config): c1, c2, c3, c3_f64, c3_learned, c3_nd, c3_ic0 / c2_ic0 (IC(0)-PCG, the
context preconditioner, on the same systems), c4 (64
right-hand sides on a 2048^2 grid and 200 random geometric graphs in one
launch sequence), c4a, c4b.

Usage: python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--config north_star]
"""
import argparse
import ctypes
import importlib
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "parallel-randomized-cholesky_amd")
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_package():
    spec = importlib.util.spec_from_file_location("rchol_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rchol_amd"] = mod
    spec.loader.exec_module(mod)
    return mod, importlib.import_module("rchol_amd.workloads")


def device_sync():
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipDeviceSynchronize()
    except OSError:
        pass


C3 = dict(dims=(256, 256, 256), kind="laplacian", weights="loguniform", rhs="normal_projected")
CONFIGS = {
    "c1": dict(workload="grid2d_64x64_unit_nnzsort", dims=(64, 64), kind="laplacian", weights="unit", prec=64,
               rhs="normal_projected"),
    "c2": dict(workload="sddm3d_128_ground_nnzsort", dims=(128, 128, 128), kind="sddm", prec=64, rhs="normal"),
    "c3": dict(C3, workload="grid3d_256_loguniform_nnzsort_f32factor", prec=32),
    "c3_f64": dict(C3, workload="grid3d_256_loguniform_nnzsort_f64", prec=64),
    "c3_learned": dict(C3, workload="grid3d_256_loguniform_learned_f32factor", prec=32, order="learned"),
    "c3_nd": dict(C3, workload="grid3d_256_loguniform_nd_f32factor", prec=32, order="nd"),
    "c3_ic0": dict(C3, workload="grid3d_256_loguniform_nnzsort_ic0_f64", prec=64, factor="ic0"),
    "c2_ic0": dict(workload="sddm3d_128_ground_nnzsort_ic0", dims=(128, 128, 128), kind="sddm", prec=64, rhs="normal",
                   factor="ic0"),
    "c4a": dict(workload="grid2d_2048_uniform_0.5_2_nnzsort_f32factor_64rhs", dims=(2048, 2048), kind="laplacian",
                weights="uniform_0.5_2", prec=32, rhs="normal_projected", nrhs=64, wseed=11),
    "c4b": dict(workload="rgg_batch_200_n10k-200k_nnzsort_f32factor", kind="rgg_batch", nsys=200, prec=32,
                rhs="normal_projected"),
}
COMPOSITE = {
    "north_star": dict(workload="grid3d_256_loguniform_learned_f32factor_vs_nnzsort_f64",
                       target="c3_learned", compare="c3_f64"),
    "c4": dict(workload="grid2d_2048_64rhs_plus_rgg_batch_200_f32factor_one_launch_sequence",
               members=("c4a", "c4b")),
}


def rgg_batch_sizes(nsys):
    """n_i = 10,000 + floor(i * 190,000 / 199) for i < 200."""
    return [10_000 + (i * 190_000) // max(nsys - 1, 1) for i in range(nsys)]


def make_batch(wl, cfg, count=None):
    sizes = rgg_batch_sizes(cfg["nsys"])[:count]
    return [wl.random_geometric_graph(m, seed=1000 + i) for i, m in enumerate(sizes)]


def traffic_of(workload, group):
    """Bytes per launch leaving the L2 for the kernel group, from the
    committed rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes of this workload
    (profiles/traffic_<workload>.json, scaled by the copy-kernel calibration);
    None when no such measurement is committed."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    g = json.load(open(path)).get("groups", {}).get(group)
    return g["traffic_bytes"] if g else None


def make_system(wl, cfg):
    if cfg["kind"] == "sddm":
        return wl.grid_sddm(cfg["dims"], 1e-2)
    return wl.grid_laplacian(cfg["dims"], cfg.get("weights", "unit"), seed=cfg.get("wseed", 7))


# ---------------------------------------------------------------- one GPU configuration
class Run:
    """One configuration on its own handle: input, ordering, resident RHS,
    warm-up, autotune, timed steps, kernel timings."""

    def __init__(self, rc, wl, name, cfg, args, system=None, systems=None):
        self.rc, self.wl, self.name, self.cfg, self.args = rc, wl, name, cfg, args
        self.batch = cfg["kind"] == "rgg_batch"
        self.nrhs = cfg.get("nrhs", 1)
        s = rc.Solver(device=0, seed=1, precision=cfg["prec"])
        t = time.perf_counter()
        if self.batch:
            s.set_batch(systems)
        else:
            n, r, c, v = system
            s.set_triplets(n, r, c, v, kind=rc.INPUT_SDDM if cfg["kind"] == "sddm" else rc.INPUT_LAPLACIAN)
        self.input_s = time.perf_counter() - t
        order = cfg.get("order", "nnz-sort")
        t = time.perf_counter()
        if order == "learned":
            sys.path.insert(0, PKG_DIR)
            tr = importlib.import_module("rchol_amd.gnn_train")
            w = tr.load_weights()
            if w is None:
                log("[bench] no trained weights: learned ordering falls back to nnz-sort")
                order = "nnz-sort (learned fallback)"
            else:
                s.set_gnn_weights(w)
            s.order(rc.ORDER_LEARNED, 1)
        elif order == "nd":
            s.order(rc.ORDER_ND, 1)
        else:
            s.order(rc.ORDER_NNZ_SORT, 1)
        self.order_s = time.perf_counter() - t
        self.order_name = order
        # within-level row order of the sweeps: the recursive bisection of the
        # graph (computed once per matrix on the host, like the ordering); a
        # batch keeps the input order (measured neutral: every system's rows
        # fit in the L2, and the per-system bisections take seconds)
        self.locality = "input"
        self.locality_s = 0.0
        if not self.batch and args.locality == "bisection":
            s.set_locality(rc.LOCALITY_BISECTION)
            self.locality = "bisection"
            self.locality_s = s.locality()[1]
        self.rhs_kind = rc.RHS_NORMAL if cfg["rhs"] == "normal" else rc.RHS_NORMAL_PROJECTED
        self.b = s.make_rhs(self.rhs_kind, 42)
        self.B = (np.stack([s.make_rhs(self.rhs_kind, 42 + q) for q in range(self.nrhs)], axis=1)
                  if self.nrhs > 1 else None)
        self.s = s
        log(f"[bench] {name}: input {self.input_s:.1f}s, ordering ({order}) {self.order_s:.2f}s, "
            f"locality ({self.locality}) {self.locality_s:.2f}s")

    def set_rhs(self):
        if self.B is not None:
            self.s.set_rhs_multi(self.B)              # the right-hand sides stay in HBM
        else:
            self.s.set_rhs(self.b)

    def summarize(self, sts):
        worst = max(sts, key=lambda d: d["iterations"])
        return dict(worst, iterations=max(d["iterations"] for d in sts),
                    rel_res_true=max(d["rel_res_true"] for d in sts),
                    converged=int(all(d["converged"] for d in sts)))

    def solve(self):
        if self.B is not None:
            return self.summarize(self.s.solve_multi_resident(tol=self.args.tol))
        return self.s.solve_resident(self.args.tol)

    def factor(self):
        return self.s.factor_ic0() if self.cfg.get("factor") == "ic0" else self.s.factor()

    def warmup(self, steps):
        self.fst = self.factor()
        self.set_rhs()
        for w in range(max(steps, 1)):
            t = time.perf_counter()
            self.fst = self.factor()
            sst = self.solve()
            log(f"[bench] {self.name} warmup {w}: {time.perf_counter() - t:.3f}s iters={sst['iterations']} "
                f"factor={self.fst['factor_ms']:.1f}ms assemble={self.fst['assemble_ms']:.1f}ms pcg={sst['pcg_ms']:.1f}ms")
        self.tune = None
        if not self.args.no_autotune:
            t = time.perf_counter()
            if self.B is not None:
                self.s.set_rhs(self.b)                 # the tuner times the single-RHS sweeps
            self.tune = self.s.autotune(reps=3)
            self.tune["autotune_s"] = time.perf_counter() - t
            self.set_rhs()
            log(f"[bench] {self.name} autotune: tail limit {self.tune['tail_max_rows']} grid "
                f"{self.tune['sweep_blocks']} sweeps {self.tune['sweeps_ms_before']:.3f} -> "
                f"{self.tune['sweeps_ms_after']:.3f} ms ({self.tune['autotune_s']:.2f}s)")

    def timed(self, steps, label=""):
        device_sync()
        times, fms, ams, pms, iters, res = [], [], [], [], [], []
        t_all = time.perf_counter()
        for k in range(steps):
            t = time.perf_counter()
            fst = self.factor()
            sst = self.solve()
            times.append(time.perf_counter() - t)
            fms.append(fst["factor_ms"]); ams.append(fst["assemble_ms"]); pms.append(sst["pcg_ms"])
            iters.append(sst["iterations"]); res.append(sst["rel_res_true"])
            log(f"[bench] {self.name}{label} step {k}: {times[-1]:.3f}s iters={sst['iterations']} "
                f"rel_res={sst['rel_res_true']:.2e}")
        device_sync()
        self.fst, self.sst = fst, sst
        return {"times": times, "mean_s": float(np.mean(times)), "median_s": float(np.median(times)),
                "wall_s": time.perf_counter() - t_all, "factor_ms": float(np.mean(fms)),
                "assemble_ms": float(np.mean(ams)), "pcg_ms": float(np.mean(pms)), "iterations": iters,
                "rel_res_true": res[-1], "nnz_factor": fst["nnz_factor"], "fills": fst["fills"],
                "nlevels": fst["nlevels"], "max_degree": fst["max_degree"]}

    def equal_accuracy(self, steps):
        """The same steps with the true-residual stop rule ||b - A x|| / ||b|| <= tol."""
        rc = self.rc
        self.s.set_solve_options(stop_rule=rc.STOP_TRUE_RESIDUAL)
        try:
            r = self.timed(steps, label=" (true-residual rule)")
        finally:
            self.s.set_solve_options(stop_rule=rc.STOP_REFERENCE)
        return {"mean_s": r["mean_s"], "iterations": r["iterations"][-1], "rel_res_true": r["rel_res_true"]}

    def kernels(self, reps, iters):
        rc, s = self.rc, self.s
        kern = {}
        if self.B is not None:
            X = s.get_solution_multi()          # the multi-RHS timings overwrite the solve's vectors
            ids = (("spmv", rc.KERNEL_MULTI_SPMM), ("trsv_fwd", rc.KERNEL_MULTI_TRSV_FWD),
                   ("trsv_bwd", rc.KERNEL_MULTI_TRSV_BWD))
        else:
            X = None
            ids = (("spmv", rc.KERNEL_SPMV), ("trsv_fwd", rc.KERNEL_TRSV_FWD), ("trsv_bwd", rc.KERNEL_TRSV_BWD))
        for name, kid in ids:
            ms, by = s.time_kernel(kid, reps=reps)
            kern[name] = {"ms": ms, "bytes": by, "GBps": by / ms / 1e6}
            log(f"[bench] {self.name} kernel {name}: {ms:.4f} ms  {by / 1e6:.1f} MB  {by / ms / 1e6:.1f} GB/s")
        kern["factor"] = {"ms": self.fst["factor_ms"], "bytes": self.fst["bytes_factor"],
                          "GBps": self.fst["bytes_factor"] / self.fst["factor_ms"] / 1e6}
        per_solve = {k: kern[k]["ms"] * (iters + 1) for k in ("spmv", "trsv_fwd", "trsv_bwd")}
        per_solve["factor"] = self.fst["factor_ms"]
        dom = max(per_solve, key=per_solve.get)
        multi = " (k right-hand sides)" if self.B is not None else ""
        kname = {"spmv": ("k_mspmv" if self.B is not None else "k_spmv_plain") + multi,
                 "trsv_fwd": ("k_msweep<fwd>" if self.B is not None else
                              "forward sweep (k_sweep_bulk + k_tail_prep + k_tail_solve)") + multi,
                 "trsv_bwd": ("k_msweep<bwd>" if self.B is not None else
                              "backward sweep (k_tail_solve + k_sweep_bulk)") + multi,
                 "factor": "k_factor_round (all rounds)"}[dom]
        return kern, dom, kname, X

    def close(self):
        self.s.close()


def copy_calibration(rc, s, reps):
    ms, by = s.time_kernel(rc.KERNEL_COPY, reps=max(3, reps // 2))
    log(f"[bench] kernel copy (calibration): {ms:.4f} ms  {by / ms / 1e6:.1f} GB/s")
    return {"ms": ms, "bytes": by, "GBps": by / ms / 1e6}


def roofline(workload, kern, dom, kname):
    a = kern[dom]["GBps"]
    return {"bound": "hbm", "kernel": kname, "achieved": a, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": a / PEAK_HBM_GBS, "traffic": traffic_of(workload, dom)}


# ---------------------------------------------------------------- CPU baseline
def cpu_threads():
    n = os.cpu_count() or 1
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(wl, cfg, args, dims=None, nrhs=1):
    """The multi-threaded CPU restatement (oracle/rchol_omp.c: dependency-driven
    OpenMP factorization, level-scheduled OpenMP PCG) on the plain method
    (nnz-sort, FP64) of the same generator: one full-size run on all host
    threads, and a thread sweep 1, 2, 4, ... on a reduced size (median of 5)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    threads = cpu_threads()
    sc = dict(cfg, dims=dims or cfg["dims"])
    n, r, c, v = make_system(wl, sc)
    kind = "sddm" if cfg["kind"] == "sddm" else "laplacian"
    fac, sys_, perm, appended = orc.prepare_system(n, r, c, v, kind=kind, order="nnz-sort", order_seed=1)
    b = np.array([orc.normal(42, i) for i in range(n)]) if n <= 300_000 else orc.normal_vector(42, n)
    if kind != "sddm":
        b -= b.mean()
    log(f"[bench] cpu: {n} rows prepared, factor on {threads} threads ...")
    t0 = time.perf_counter()
    f = orc.factor_omp(fac, seed=1, threads=threads)
    if appended:
        f = orc.factor_trim_last(f)
    t1 = time.perf_counter()
    log(f"[bench] cpu: factor {t1 - t0:.2f}s, PCG ...")
    _, st = orc.pcg_omp(sys_, f, b[perm], tol=args.tol, threads=threads)
    t2 = time.perf_counter()
    log(f"[bench] cpu: PCG {t2 - t1:.2f}s ({st['iterations']} iterations)")
    out = {"value": (t1 - t0) + nrhs * (t2 - t1), "unit": "s", "cores": threads, "kind": "port",
           "nproc": os.cpu_count(), "factor_s": t1 - t0, "pcg_s": t2 - t1, "iterations": st["iterations"],
           "rel_res_true": st["rel_res_true"],
           "sample": (f"CPU restatement (OpenMP, {threads} threads) of the plain method (nnz-sort, FP64) on "
                      f"{'x'.join(map(str, sc['dims']))} ({n} rows): factor {t1 - t0:.2f}s + PCG {t2 - t1:.2f}s, "
                      f"{st['iterations']} iterations")}
    if nrhs > 1:
        out["sample"] += f"; PCG time x{nrhs} for the {nrhs} right-hand sides"
    if not args.no_thread_sweep:
        sd = tuple(max(2, d // 4) for d in sc["dims"]) if len(sc["dims"]) == 3 else tuple(max(2, d // 8) for d in sc["dims"])
        sc2 = dict(cfg, dims=sd)
        n2, r2, c2, v2 = make_system(wl, sc2)
        fac2, sys2, perm2, app2 = orc.prepare_system(n2, r2, c2, v2, kind=kind, order="nnz-sort", order_seed=1)
        b2 = orc.normal_vector(42, n2)
        if kind != "sddm":
            b2 -= b2.mean()
        sweep = []
        t = 1
        while True:
            fs, ps = [], []
            for _ in range(5):
                a0 = time.perf_counter()
                f2 = orc.factor_omp(fac2, seed=1, threads=t)
                if app2:
                    f2 = orc.factor_trim_last(f2)
                a1 = time.perf_counter()
                orc.pcg_omp(sys2, f2, b2[perm2], tol=args.tol, threads=t)
                a2 = time.perf_counter()
                fs.append(a1 - a0); ps.append(a2 - a1)
            sweep.append({"threads": t, "factor_s": float(np.median(fs)), "pcg_s": float(np.median(ps))})
            log(f"[bench] cpu sweep: {t} threads factor {sweep[-1]['factor_s']:.3f}s PCG {sweep[-1]['pcg_s']:.3f}s")
            if t >= threads:
                break
            t = min(2 * t, threads)
        out["thread_sweep"] = {"workload": f"{'x'.join(map(str, sd))} ({n2} rows), same generator",
                               "median_of": 5, "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"), "runs": sweep}
    return out


def component_means(n, rows, cols, b):
    """Mean of b over each vertex's connected component (RHS projection)."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components
    g = sp.coo_matrix((np.ones(len(rows)), (rows, cols)), shape=(n, n))
    nc, lab = connected_components(g, directed=False)
    s = np.bincount(lab, weights=b, minlength=nc)
    c = np.bincount(lab, minlength=nc)
    return (s / c)[lab]


def cpu_baseline_batch(wl, cfg, args, max_n):
    """The CPU restatement on the first systems of the batch, each solved
    alone on all host threads, until max_n rows are covered; time scaled by
    total rows / sampled rows."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    threads = cpu_threads()
    sizes = rgg_batch_sizes(cfg["nsys"])
    total = sum(sizes)
    t_f = t_p = 0.0
    rows = 0
    its = []
    i = 0
    while i < len(sizes) and rows < max_n:
        n, r, c, v = wl.random_geometric_graph(sizes[i], seed=1000 + i)
        fac, sys_, perm, _ = orc.prepare_system(n, r, c, v, kind="laplacian", order="nnz-sort", order_seed=1)
        b = orc.normal_vector(42, n)
        b -= component_means(n, r, c, b)          # consistent per connected component, as on the GPU
        t0 = time.perf_counter()
        f = orc.factor_omp(fac, seed=1, threads=threads)
        t1 = time.perf_counter()
        _, st = orc.pcg_omp(sys_, f, b[perm], tol=args.tol, threads=threads)
        t2 = time.perf_counter()
        t_f += t1 - t0
        t_p += t2 - t1
        rows += n
        its.append(st["iterations"])
        i += 1
    scale = total / rows
    sample = (f"CPU restatement (OpenMP, {threads} threads, FP64) of the first {i} systems ({rows} rows) one by "
              f"one: factor {t_f:.2f}s + PCG {t_p:.2f}s, iterations {min(its)}-{max(its)}; scaled x{scale:.1f} "
              f"to {total} rows")
    return {"value": (t_f + t_p) * scale, "unit": "s", "cores": threads, "kind": "port", "nproc": os.cpu_count(),
            "sample": sample, "factor_s": t_f, "pcg_s": t_p}


# ---------------------------------------------------------------- drivers
def error_vs_fp64(rc, run, system, args):
    """||x - x_ref||_A / ||x_ref||_A of the FP32-factor solution, x_ref from an
    FP64 factor (nnz-sort) and FP64 PCG to 1e-12 of the same system."""
    t = time.perf_counter()
    s64 = rc.Solver(device=0, seed=1, precision=64, maxit=3000)
    n, r, c, v = system
    s64.set_triplets(n, r, c, v, kind=rc.INPUT_SDDM if run.cfg["kind"] == "sddm" else rc.INPUT_LAPLACIAN)
    s64.order(rc.ORDER_NNZ_SORT, 1)
    s64.factor()
    x_ref, st64 = s64.solve(run.b, tol=1e-12)
    s64.close()
    run.s.set_reference_solution(x_ref)
    _, ste = run.s.solve(run.b, tol=args.tol)
    run.s.set_reference_solution(None)
    log(f"[bench] {run.name} error vs FP64 reference: {ste['err_vs_oracle']:.3e} ({time.perf_counter() - t:.1f}s)")
    return {"err_A_norm": ste["err_vs_oracle"], "iterations": ste["iterations"],
            "reference": (f"FP64 factor (nnz-sort) + FP64 PCG to 1e-12 on the GPU: {st64['iterations']} iterations, "
                          f"true rel-res {st64['rel_res_true']:.1e}")}


def base_line(args, metric, value, dtype, config):
    return {"metric": metric, "value": value, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": value * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": dtype, "data": "synthetic", "config": config}


def bench_north_star(rc, wl, args):
    comp = COMPOSITE["north_star"]
    cfg_t, cfg_c = CONFIGS[comp["target"]], CONFIGS[comp["compare"]]
    t = time.perf_counter()
    system = make_system(wl, cfg_t)
    log(f"[bench] generated n={system[0]} triplets={len(system[1])} in {time.perf_counter() - t:.1f}s")
    out = {}
    runs = {}
    for key, name, cfg in (("plain", comp["compare"], cfg_c), ("learned", comp["target"], cfg_t)):
        run = Run(rc, wl, name, cfg, args, system=system)
        run.warmup(args.warmup)
        res = run.timed(args.steps)
        res["equal_accuracy"] = run.equal_accuracy(min(2, args.steps))
        res["order_s"] = run.order_s
        res["input_s"] = run.input_s
        res["locality"] = run.locality
        res["locality_s"] = run.locality_s
        res["autotune"] = run.tune
        runs[key] = run
        out[key] = res
    learned, plain = runs["learned"], runs["plain"]
    kern, dom, kname, _ = learned.kernels(args.kernel_reps, np.mean(out["learned"]["iterations"]))
    kern["copy_1GiB"] = copy_calibration(rc, learned.s, args.kernel_reps)
    out["learned"]["kernels"] = kern
    pk, pdom, pkname, _ = plain.kernels(args.kernel_reps, np.mean(out["plain"]["iterations"]))
    out["plain"]["kernels"] = pk
    out["learned"]["error_vs_fp64"] = error_vs_fp64(rc, learned, system, args)
    speed = out["plain"]["mean_s"] / out["learned"]["mean_s"]
    speed_eq = out["plain"]["equal_accuracy"]["mean_s"] / out["learned"]["equal_accuracy"]["mean_s"]
    log(f"[bench] learned/FP32 {out['learned']['mean_s']:.3f}s vs plain FP64 {out['plain']['mean_s']:.3f}s: "
        f"{speed:.2f}x (true-residual rule: {speed_eq:.2f}x)")
    cfg_line = {"workload": comp["workload"], "n": int(learned.s.size()[0]), "nnz": int(learned.s.size()[1]),
                "target": {"order": learned.order_name, "factor_precision": 32},
                "compare": {"order": "nnz-sort", "factor_precision": 64}, "tol": args.tol,
                "stop_rule": "reference sqrt(r'z/r0'z0)"}
    res = base_line(args, "factor+PCG time to rel-res 1e-6 (s); learned/FP32 vs plain FP64; HBM GB/s",
                    out["learned"]["mean_s"], "f32-factor/f64-pcg", cfg_line)
    res["speedup_vs_plain_fp64"] = speed
    res["speedup_vs_plain_fp64_true_residual_rule"] = speed_eq
    res["plain_fp64_s"] = out["plain"]["mean_s"]
    res["learned_ordering_s"] = learned.order_s
    res["roofline"] = roofline(CONFIGS[comp["target"]]["workload"], kern, dom, kname)
    res["detail"] = out
    for r in runs.values():
        r.close()
    return res, cfg_c


def bench_c4(rc, wl, args):
    """64 right-hand sides on the 2048^2 grid and the 200-graph batch, solved
    together by rc_solve_group (one captured graph per two iterations)."""
    ca, cb = CONFIGS["c4a"], CONFIGS["c4b"]
    t = time.perf_counter()
    grid = make_system(wl, ca)
    systems = make_batch(wl, cb)
    log(f"[bench] generated grid n={grid[0]} and {len(systems)} graphs (n={sum(q[0] for q in systems)}) "
        f"in {time.perf_counter() - t:.1f}s")
    ra = Run(rc, wl, "c4a", ca, args, system=grid)
    rb = Run(rc, wl, "c4b", cb, args, systems=systems)
    del systems
    for r in (ra, rb):
        r.warmup(1)

    def step():
        fa = ra.s.factor()
        fb = rb.s.factor()
        summ = rc.solve_group([ra.s, rb.s], tol=args.tol)
        return fa, fb, summ

    for w in range(max(args.warmup, 1)):
        t = time.perf_counter()
        step()
        log(f"[bench] c4 warmup {w}: {time.perf_counter() - t:.3f}s")
    device_sync()
    times, its = [], []
    t_all = time.perf_counter()
    for k in range(args.steps):
        t = time.perf_counter()
        fa, fb, summ = step()
        times.append(time.perf_counter() - t)
        its.append([summ[0]["iterations"], summ[1]["iterations"]])
        log(f"[bench] c4 step {k}: {times[-1]:.3f}s iters grid={summ[0]['iterations']} batch={summ[1]['iterations']}")
    device_sync()
    wall = time.perf_counter() - t_all
    per_rhs = ra.s.multi_stats()
    per_sys = rb.s.batch_stats()
    ra.fst, rb.fst = fa, fb
    kern, dom, kname, _ = ra.kernels(args.kernel_reps, np.mean([i[0] for i in its]))
    kern["copy_1GiB"] = copy_calibration(rc, ra.s, args.kernel_reps)
    kb, domb, knameb, _ = rb.kernels(args.kernel_reps, np.mean([i[1] for i in its]))
    cfg_line = {"workload": COMPOSITE["c4"]["workload"], "grid_n": int(grid[0]), "nrhs": 64,
                "systems": rb.s.batch()[0], "batch_n": int(rb.s.size()[0]), "factor_precision": 32,
                "order": "nnz-sort", "tol": args.tol, "stop_rule": "reference sqrt(r'z/r0'z0)"}
    res = base_line(args, "factor+PCG time to rel-res 1e-6 (s); learned/FP32 vs plain FP64; HBM GB/s",
                    float(np.mean(times)), "f32-factor/f64-pcg", cfg_line)
    res["roofline"] = roofline(ca["workload"], kern, dom, kname)
    res["detail"] = {"wall_s_timed_loop": wall, "iterations": its,
                     "grid": {"factor_ms": fa["factor_ms"], "assemble_ms": fa["assemble_ms"],
                              "iterations_per_rhs": [d["iterations"] for d in per_rhs],
                              "rel_res_true_max": max(d["rel_res_true"] for d in per_rhs), "kernels": kern},
                     "batch": {"factor_ms": fb["factor_ms"], "assemble_ms": fb["assemble_ms"],
                               "iterations_max": max(d["iterations"] for d in per_sys),
                               "rel_res_true_max": max(d["rel_res_true"] for d in per_sys), "kernels": kb},
                     "autotune": {"grid": ra.tune, "batch": rb.tune}}
    for r in (ra, rb):
        r.close()
    return res


def bench_single(rc, wl, args, name):
    cfg = dict(CONFIGS[name])
    if args.precision:
        cfg["prec"] = args.precision
    t = time.perf_counter()
    batch = cfg["kind"] == "rgg_batch"
    system = systems = None
    if batch:
        systems = make_batch(wl, cfg)
        log(f"[bench] generated {len(systems)} systems, n={sum(q[0] for q in systems)} in {time.perf_counter() - t:.1f}s")
    else:
        system = make_system(wl, cfg)
        log(f"[bench] generated n={system[0]} triplets={len(system[1])} in {time.perf_counter() - t:.1f}s")
    run = Run(rc, wl, name, cfg, args, system=system, systems=systems)
    del systems
    run.warmup(args.warmup)
    res_t = run.timed(args.steps)
    kern, dom, kname, _ = run.kernels(args.kernel_reps, np.mean(res_t["iterations"]))
    kern["copy_1GiB"] = copy_calibration(rc, run.s, args.kernel_reps)
    err64 = None
    if cfg["prec"] == 32 and run.B is None and not batch:
        err64 = error_vs_fp64(rc, run, system, args)
    cfg_line = {"workload": cfg["workload"], "n": int(run.s.size()[0]), "nnz": int(run.s.size()[1]),
                "factor_precision": cfg["prec"], "order": run.order_name, "nrhs": run.nrhs,
                "systems": run.s.batch()[0], "tol": args.tol, "stop_rule": "reference sqrt(r'z/r0'z0)"}
    res = base_line(args, "factor+PCG time to rel-res 1e-6 (s); learned/FP32 vs plain FP64; HBM GB/s",
                    res_t["mean_s"], "f64" if cfg["prec"] == 64 else "f32-factor/f64-pcg", cfg_line)
    res["roofline"] = roofline(cfg["workload"], kern, dom, kname)
    res_t.update({"kernels": kern, "autotune": run.tune, "error_vs_fp64": err64, "order_s": run.order_s,
                  "input_s": run.input_s, "locality": run.locality, "locality_s": run.locality_s})
    res["detail"] = res_t
    run.close()
    return res, cfg


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="north_star", choices=sorted(CONFIGS) + sorted(COMPOSITE))
    ap.add_argument("--precision", type=int, default=0, help="override factor precision (32/64)")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-thread-sweep", action="store_true")
    ap.add_argument("--cpu-max-n", type=int, default=2_200_000, help="rows of the CPU leg of the batched workloads")
    ap.add_argument("--kernel-reps", type=int, default=10)
    ap.add_argument("--locality", default="bisection", choices=["input", "bisection"],
                    help="within-level row order of the sweeps and SpMV tiles (bisection: recursive bisection "
                         "of the graph, computed once per matrix on the host, outside the timed steps like the "
                         "ordering; input: the input numbering)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the per-structure sweep launch tuning done once before the timed steps")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU (the library has no multi-GPU path); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    rc, wl = load_package()
    log(f"[bench] config {args.config}")
    if args.config == "north_star":
        res, cpu_cfg = bench_north_star(rc, wl, args)
        cpu = None if args.no_cpu else (lambda: cpu_baseline(wl, cpu_cfg, args))
    elif args.config == "c4":
        res = bench_c4(rc, wl, args)
        cpu = None if args.no_cpu else (lambda: {
            "grid_64rhs": cpu_baseline(wl, CONFIGS["c4a"], args, nrhs=64),
            "batch": cpu_baseline_batch(wl, CONFIGS["c4b"], args, args.cpu_max_n)})
    else:
        res, cfg = bench_single(rc, wl, args, args.config)
        if args.no_cpu:
            cpu = None
        elif cfg["kind"] == "rgg_batch":
            cpu = lambda: cpu_baseline_batch(wl, cfg, args, args.cpu_max_n)
        else:
            cpu = lambda: cpu_baseline(wl, dict(cfg, prec=64), args, nrhs=cfg.get("nrhs", 1))
    if cpu is not None:
        log("[bench] CPU baseline ...")
        c = cpu()
        if "value" not in c:                      # composite: sum of the parts
            parts = c
            c = {"value": sum(p["value"] for p in parts.values()), "unit": "s",
                 "cores": max(p["cores"] for p in parts.values()), "kind": "port",
                 "sample": "; ".join(f"{k}: {p['sample']}" for k, p in parts.items()), "parts": parts}
        res["cpu_baseline"] = c
        log(f"[bench] cpu: {c['sample']}")
    else:
        res["cpu_baseline"] = None
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
