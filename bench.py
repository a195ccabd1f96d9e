#!/usr/bin/env python3
"""Benchmark: randomized-Cholesky factor + PCG to relative residual 1e-6 on one
MI355X, with the CPU oracle timed beside it.

One step = rc_factor (elimination + assembly of the resident system) +
rc_solve_resident (PCG to tol with the RHS already in HBM).  Both calls return
after synchronising the library's stream; the timed loop is additionally
bracketed by hipDeviceSynchronize.

Default workload: 3D 7-point Poisson SDDM 128^3 with
diagonal = degree + 1e-2, converted to a Laplacian with an appended ground
vertex, nnz-sorted ordering (seed 1), FP64 factor and PCG, one N(0,1) RHS
(seed 42), tol 1e-6.

Usage: python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--config c2]
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


CONFIGS = {
    "c1": dict(workload="grid2d_64x64_unit_nnzsort", dims=(64, 64), kind="laplacian", weights="unit", prec=64,
               rhs="normal_projected"),
    "c2": dict(workload="sddm3d_128_ground_nnzsort", dims=(128, 128, 128), kind="sddm", prec=64, rhs="normal"),
    "c3": dict(workload="grid3d_256_loguniform_nnzsort_f32factor", dims=(256, 256, 256), kind="laplacian",
               weights="loguniform", prec=32, rhs="normal_projected"),
    "c3_f64": dict(workload="grid3d_256_loguniform_nnzsort_f64", dims=(256, 256, 256), kind="laplacian",
                   weights="loguniform", prec=64, rhs="normal_projected"),
    "c3_learned": dict(workload="grid3d_256_loguniform_learned_f32factor", dims=(256, 256, 256), kind="laplacian",
                       weights="loguniform", prec=32, rhs="normal_projected", order="learned"),
    "c4a": dict(workload="grid2d_2048_uniform_0.5_2_nnzsort_f32factor_64rhs", dims=(2048, 2048), kind="laplacian",
                weights="uniform_0.5_2", prec=32, rhs="normal_projected", nrhs=64, wseed=11),
    "c4b": dict(workload="rgg_batch_200_n10k-200k_nnzsort_f32factor", kind="rgg_batch", nsys=200, prec=32,
                rhs="normal_projected"),
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


def cpu_baseline_batch(wl, cfg, tol, max_n):
    """The oracle on the first systems of the batch, each solved alone, until
    max_n rows are covered; time scaled by total rows / sampled rows."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    sizes = rgg_batch_sizes(cfg["nsys"])
    total = sum(sizes)
    t_f = t_p = 0.0
    rows = 0
    its = []
    i = 0
    while i < len(sizes) and rows < max_n:
        n, r, c, v = wl.random_geometric_graph(sizes[i], seed=1000 + i)
        fac, sys_, perm, _ = orc.prepare_system(n, r, c, v, kind="laplacian", order="nnz-sort", order_seed=1)
        b = np.array([orc.normal(42, j) for j in range(n)])
        b -= b.mean()
        t0 = time.perf_counter()
        f = orc.factor(fac, seed=1)
        t1 = time.perf_counter()
        _, st = orc.pcg(sys_, f, b[perm], tol=tol, precision=cfg["prec"])
        t2 = time.perf_counter()
        t_f += t1 - t0
        t_p += t2 - t1
        rows += n
        its.append(st["iterations"])
        i += 1
    scale = total / rows
    sample = (f"oracle (1 thread) factor+PCG of the first {i} systems ({rows} rows) one by one: factor {t_f:.2f}s "
              f"+ PCG {t_p:.2f}s, iterations {min(its)}-{max(its)}; scaled x{scale:.1f} to {total} rows")
    return {"value": (t_f + t_p) * scale, "unit": "s", "cores": 1, "kind": "port", "sample": sample,
            "factor_s": t_f, "pcg_s": t_p}


def cpu_baseline(wl, cfg, b_full, tol, max_n):
    """The oracle (single-threaded C port of the reference algorithm) on the
    same workload, or on the same generator at a reduced size when the full
    size exceeds max_n rows (time scaled linearly in n; stated in `sample`)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    dims = cfg["dims"]
    full_n = int(np.prod(dims))
    scale = 1.0
    sdims = dims
    while int(np.prod(sdims)) > max_n:
        sdims = tuple(max(2, d // 2) for d in sdims)
    sc = dict(cfg, dims=sdims)
    n, r, c, v = make_system(wl, sc)
    kind = "sddm" if cfg["kind"] == "sddm" else "laplacian"
    fac, sys_, perm, appended = orc.prepare_system(n, r, c, v, kind=kind, order="nnz-sort", order_seed=1)
    if sdims == dims:
        b = b_full
    else:
        b = np.array([orc.normal(42, i) for i in range(n)])
        if kind != "sddm":
            b -= b.mean()
        scale = full_n / n
    t0 = time.perf_counter()
    f = orc.factor(fac, seed=1)
    if appended:
        f = orc.factor_trim_last(f)
    t1 = time.perf_counter()
    x, st = orc.pcg(sys_, f, b[perm], tol=tol, precision=cfg["prec"])
    t2 = time.perf_counter()
    nrhs = cfg.get("nrhs", 1)
    sample = (f"oracle (1 thread) factor+PCG on {'x'.join(map(str, sdims))} "
              f"({n} rows, nnz-sort order): factor {t1 - t0:.2f}s + PCG {t2 - t1:.2f}s, {st['iterations']} iterations")
    if nrhs > 1:
        sample += f"; one right-hand side solved, PCG time x{nrhs} for the {nrhs} right-hand sides"
    if scale != 1.0:
        sample += f"; scaled x{scale:.1f} to {full_n} rows"
    return {"value": ((t1 - t0) + nrhs * (t2 - t1)) * scale, "unit": "s", "cores": 1, "kind": "port",
            "sample": sample, "iterations": st["iterations"], "factor_s": t1 - t0, "pcg_s": t2 - t1}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--precision", type=int, default=0, help="override factor precision (32/64)")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-max-n", type=int, default=2_200_000)
    ap.add_argument("--kernel-reps", type=int, default=10)
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the per-structure sweep launch tuning done once before the timed steps")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU (the library has no multi-GPU path); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    rc, wl = load_package()
    cfg = dict(CONFIGS[args.config])
    if args.precision:
        cfg["prec"] = args.precision
    log(f"[bench] config {args.config}: generating {cfg['workload']}")
    t = time.perf_counter()
    batch = cfg["kind"] == "rgg_batch"
    if batch:
        systems = make_batch(wl, cfg)
        log(f"[bench] generated {len(systems)} systems, n={sum(q[0] for q in systems)} in {time.perf_counter() - t:.1f}s")
    else:
        n, r, c, v = make_system(wl, cfg)
        log(f"[bench] generated n={n} triplets={len(r)} in {time.perf_counter() - t:.1f}s")

    s = rc.Solver(device=0, seed=1, precision=cfg["prec"])
    t = time.perf_counter()
    if batch:
        s.set_batch(systems)
        del systems
    else:
        s.set_triplets(n, r, c, v, kind=rc.INPUT_SDDM if cfg["kind"] == "sddm" else rc.INPUT_LAPLACIAN)
    order_name = cfg.get("order", "nnz-sort")
    if order_name == "learned":
        sys.path.insert(0, PKG_DIR)
        tr = importlib.import_module("rchol_amd.gnn_train")
        w = tr.load_weights()
        if w is None:
            log("[bench] no trained weights: learned ordering falls back to nnz-sort")
            order_name = "nnz-sort (learned fallback)"
        else:
            s.set_gnn_weights(w)
        t_o = time.perf_counter()
        s.order(rc.ORDER_LEARNED, 1)
        log(f"[bench] learned ordering {time.perf_counter() - t_o:.2f}s")
    else:
        s.order(rc.ORDER_NNZ_SORT, 1)
    nrhs = cfg.get("nrhs", 1)
    rhs_kind = rc.RHS_NORMAL if cfg["rhs"] == "normal" else rc.RHS_NORMAL_PROJECTED
    b = s.make_rhs(rhs_kind, 42)
    B = np.stack([s.make_rhs(rhs_kind, 42 + q) for q in range(nrhs)], axis=1) if nrhs > 1 else None
    log(f"[bench] host input path (Laplacian, ordering, RHS) {time.perf_counter() - t:.1f}s")

    # warmup: first factor also prepares and uploads the system
    fst = sst = None
    if B is not None:
        s.set_rhs_multi(B)                         # the 64 right-hand sides stay in HBM

    def solve_step():
        if B is not None:
            sts = s.solve_multi_resident(tol=args.tol)
            worst = max(sts, key=lambda d: d["iterations"])
            return dict(worst, iterations=max(d["iterations"] for d in sts),
                        rel_res_true=max(d["rel_res_true"] for d in sts))
        return s.solve_resident(args.tol)

    for w in range(max(args.warmup, 1)):
        t = time.perf_counter()
        fst = s.factor()
        s.set_rhs(b)
        sst = solve_step()
        log(f"[bench] warmup {w}: {time.perf_counter() - t:.3f}s iters={sst['iterations']} "
            f"factor={fst['factor_ms']:.1f}ms assemble={fst['assemble_ms']:.1f}ms pcg={sst['pcg_ms']:.1f}ms")

    tune = None
    if not args.no_autotune:
        t = time.perf_counter()
        tune = s.autotune(reps=3)
        tune["autotune_s"] = time.perf_counter() - t
        log(f"[bench] autotune: tail limit {tune['tail_max_rows']} grid {tune['sweep_blocks']} sweeps "
            f"{tune['sweeps_ms_before']:.3f} -> {tune['sweeps_ms_after']:.3f} ms ({tune['autotune_s']:.2f}s)")

    device_sync()
    times, fms, ams, pms, iters = [], [], [], [], []
    t_all = time.perf_counter()
    for k in range(args.steps):
        t = time.perf_counter()
        fst = s.factor()
        sst = solve_step()
        times.append(time.perf_counter() - t)
        fms.append(fst["factor_ms"]); ams.append(fst["assemble_ms"]); pms.append(sst["pcg_ms"])
        iters.append(sst["iterations"])
        log(f"[bench] step {k}: {times[-1]:.3f}s iters={sst['iterations']} rel_res={sst['rel_res_true']:.2e}")
    device_sync()
    wall = time.perf_counter() - t_all

    # per-kernel timing on the library stream (hipEvents), algorithmic bytes
    kern = {}
    for name, kid in (("spmv", rc.KERNEL_SPMV), ("trsv_fwd", rc.KERNEL_TRSV_FWD), ("trsv_bwd", rc.KERNEL_TRSV_BWD)):
        ms, by = s.time_kernel(kid, reps=args.kernel_reps)
        kern[name] = {"ms": ms, "bytes": by, "GBps": by / ms / 1e6}
        log(f"[bench] kernel {name}: {ms:.4f} ms  {by / 1e6:.1f} MB  {by / ms / 1e6:.1f} GB/s")
    ms, by = s.time_kernel(rc.KERNEL_COPY, reps=max(3, args.kernel_reps // 2))
    kern["copy_1GiB"] = {"ms": ms, "bytes": by, "GBps": by / ms / 1e6}
    log(f"[bench] kernel copy (calibration): {ms:.4f} ms  {by / ms / 1e6:.1f} GB/s")
    fac_gbs = fst["bytes_factor"] / fst["factor_ms"] / 1e6
    kern["factor"] = {"ms": fst["factor_ms"], "bytes": fst["bytes_factor"], "GBps": fac_gbs}
    per_solve = {k: kern[k]["ms"] * (np.mean(iters) + 1) for k in ("spmv", "trsv_fwd", "trsv_bwd")}
    per_solve["factor"] = fst["factor_ms"]
    dom = max(per_solve, key=per_solve.get)
    kname = {"spmv": "k_spmv_plain", "trsv_fwd": "forward sweep (k_sweep_bulk + k_tail_prep + k_tail_solve)",
             "trsv_bwd": "backward sweep (k_tail_solve + k_sweep_bulk)", "factor": "k_factor_round (all rounds)"}[dom]
    achieved = kern[dom]["GBps"]

    x = s.get_solution()
    # mixed precision: error of the FP32-factor solution against an FP64
    # oracle of the same system (FP64 factor and PCG to 1e-12; a direct solve
    # of this size is out of reach), measured outside the timed steps
    err64 = None
    if cfg["prec"] == 32 and B is None and not batch:
        t = time.perf_counter()
        s64 = rc.Solver(device=0, seed=1, precision=64, maxit=3000)
        s64.set_triplets(n, r, c, v, kind=rc.INPUT_SDDM if cfg["kind"] == "sddm" else rc.INPUT_LAPLACIAN)
        s64.order(rc.ORDER_NNZ_SORT, 1)
        s64.factor()
        x_ref, st64 = s64.solve(b, tol=1e-12)
        s64.close()
        s.set_reference_solution(x_ref)
        _, ste = s.solve(b, tol=args.tol)
        err64 = {"err_A_norm": ste["err_vs_oracle"], "iterations": ste["iterations"],
                 "oracle": (f"FP64 factor (nnz-sort) + FP64 PCG to 1e-12: {st64['iterations']} iterations, "
                            f"true rel-res {st64['rel_res_true']:.1e}")}
        log(f"[bench] error vs FP64 oracle: {ste['err_vs_oracle']:.3e} (A-norm, relative) "
            f"({time.perf_counter() - t:.1f}s)")
    res = {
        "metric": "factor+PCG time to rel-res 1e-6 (s)",
        "value": float(np.mean(times)),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(np.mean(times) * 1e3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if cfg["prec"] == 64 else "f32-factor/f64-pcg",
        "data": "synthetic",
        "config": {"workload": cfg["workload"], "n": int(s.size()[0]), "nnz": int(s.size()[1]),
                   "factor_precision": cfg["prec"], "order": order_name, "nrhs": nrhs,
                   "systems": s.batch()[0], "tol": args.tol,
                   "stop_rule": "reference sqrt(r'z/r0'z0)"},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic_of(cfg["workload"], dom)},
        "detail": {
            "wall_s_timed_loop": wall,
            "factor_ms": float(np.mean(fms)), "assemble_ms": float(np.mean(ams)), "pcg_ms": float(np.mean(pms)),
            "iterations": iters, "rel_res_true": sst["rel_res_true"], "nnz_factor": fst["nnz_factor"],
            "fills": fst["fills"], "nlevels": fst["nlevels"], "max_degree": fst["max_degree"],
            "kernels": kern,
            "autotune": tune,
            "error_vs_fp64_oracle": err64,
        },
    }
    if not args.no_cpu:
        log("[bench] CPU oracle baseline ...")
        res["cpu_baseline"] = (cpu_baseline_batch(wl, cfg, args.tol, args.cpu_max_n) if batch
                               else cpu_baseline(wl, cfg, b, args.tol, args.cpu_max_n))
        log(f"[bench] cpu: {res['cpu_baseline']['sample']}")
    else:
        res["cpu_baseline"] = None
    s.close()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
