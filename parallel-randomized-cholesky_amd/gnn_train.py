"""Trainer of the ordering network (no counterpart in the reference, whose
orderings are fixed Julia heuristics, write_graph.jl:46-78).

The network scores every vertex (csrc/device/gnn.hip); an ordering is drawn
from the Plackett-Luce distribution of the scores (sort of score + Gumbel
noise, descending = eliminated first).  Each sampled ordering is run through
the method itself on the GPU -- randomized factor + PCG to the tolerance --
and its cost
    cost = log(PCG iterations) + log(levels of the triangular solves)
           + 0.25 log(factor nnz)
is the training signal (REINFORCE with the batch mean as baseline).  On the
GPU the elimination takes one round per level and every PCG iteration walks
all levels twice, so iterations x levels is the proxy for the time.  The
network is first fitted (mean squared error, same backward kernels) to a
simple prior -- eliminate vertices of small weighted degree first, scores
-2 (log weighted degree - mean) -- and then trained on the method's costs.  The
gradient of log p(ordering) with respect to the scores has a closed form
(prefix sums over the ordering); the gradient with respect to the weights comes
from the library's backward kernels (rc_gnn_gradient).  Adam runs on the host
(16.9k parameters).

Usage: python gnn_train.py [--steps 300] [--out gnn_weights.f32]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(_HERE, "gnn_weights.f32")
HIDDEN = 64


def _pkg():
    mod = sys.modules.get("rchol_amd")
    if mod is not None:
        return mod
    import importlib.util
    spec = importlib.util.spec_from_file_location("rchol_amd", os.path.join(_HERE, "__init__.py"),
                                                  submodule_search_locations=[_HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rchol_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def init_weights(nparam: int, seed: int = 0) -> np.ndarray:
    """He-style initialisation in the library's weight layout."""
    rng = np.random.default_rng(seed)
    w = np.zeros(nparam, dtype=np.float32)
    off = 0
    for din in (2, HIDDEN, HIDDEN):
        std = np.sqrt(2.0 / (2 * din))
        w[off:off + 2 * HIDDEN * din] = rng.normal(0.0, std, 2 * HIDDEN * din)
        off += 2 * HIDDEN * din
        w[off:off + HIDDEN] = 0.01
        off += HIDDEN
    w[off:off + HIDDEN] = rng.normal(0.0, np.sqrt(1.0 / HIDDEN), HIDDEN)
    return w


def load_weights(path: str = WEIGHTS):
    if not os.path.exists(path):
        return None
    return np.fromfile(path, dtype=np.float32)


def pl_logp_grad(scores: np.ndarray, perm: np.ndarray) -> np.ndarray:
    """d log p(perm | scores) / d scores for the Plackett-Luce model with
    perm[0] chosen first: log p = sum_i s_{p_i} - log sum_{j >= i} exp(s_{p_j})."""
    s = scores.astype(np.float64)
    e = np.exp(s[perm] - s.max())
    Z = np.cumsum(e[::-1])[::-1]
    inv = np.cumsum(1.0 / Z)
    g = np.empty_like(s)
    g[perm] = 1.0 - e * inv
    return g


def cost_of(fs: dict, ss: dict) -> float:
    return float(np.log(max(ss["iterations"], 1)) + np.log(max(fs["nlevels"], 1))
                 + 0.25 * np.log(max(fs["nnz_factor"], 1)))


def prior_scores(n, rows, cols, vals, kappa=2.0):
    """-kappa (log weighted degree - mean over vertices with edges)."""
    rows = np.asarray(rows); cols = np.asarray(cols); vals = np.abs(np.asarray(vals, dtype=np.float64))
    off = rows != cols
    wdeg = np.bincount(rows[off], weights=vals[off], minlength=n)
    t = np.zeros(n)
    has = wdeg > 0
    if has.any():
        lw = np.log(wdeg[has])
        t[has] = -kappa * (lw - lw.mean())
    return t


def training_graphs(wl, small: bool = False):
    g = []
    sizes2 = [(40, 40), (56, 56)] if small else [(48, 48), (64, 64), (96, 96)]
    sizes3 = [(12, 12, 12)] if small else [(14, 14, 14), (18, 18, 18), (24, 24, 24)]
    for i, d in enumerate(sizes2):
        g.append((f"grid{d}_log", wl.grid_laplacian(d, "loguniform", seed=100 + i)))
        g.append((f"grid{d}_u", wl.grid_laplacian(d, "uniform_0.5_2", seed=200 + i)))
    for i, d in enumerate(sizes3):
        g.append((f"grid{d}_log", wl.grid_laplacian(d, "loguniform", seed=300 + i)))
    g.append(("rgg5000", wl.random_geometric_graph(5000 if small else 8000, seed=400)))
    return g


def evaluate(rc, w, graph, seed=1, precision=64, tol=1e-6):
    """Cost of the learned ordering, of the prior it starts from and of the
    nnz-sort fallback on one graph."""
    n, r, c, v = graph
    out = {}
    for name in ("nnz-sort", "prior", "learned"):
        s = rc.Solver(device=0, seed=seed, precision=precision)
        s.set_triplets(n, r, c, v)
        if name == "learned":
            s.set_gnn_weights(w)
            s.order(rc.ORDER_LEARNED, seed)
        elif name == "prior":
            g = np.random.default_rng(seed).gumbel(size=n)
            s.set_permutation(np.argsort(-(prior_scores(n, r, c, v) + g), kind="stable").astype(np.int32))
        else:
            s.order(rc.ORDER_NNZ_SORT, seed)
        fs = s.factor()
        b = s.make_rhs(rc.RHS_NORMAL_PROJECTED, 42)
        _, ss = s.solve(b, tol=tol)
        out[name] = {"iterations": ss["iterations"], "nlevels": fs["nlevels"], "nnz_factor": fs["nnz_factor"],
                     "cost": cost_of(fs, ss), "factor_ms": fs["factor_ms"], "pcg_ms": ss["pcg_ms"]}
        s.close()
    return out


def adam_step(w, dw, m, v2, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * dw
    v2 = b2 * v2 + (1 - b2) * dw * dw
    mh = m / (1 - b1 ** (step + 1))
    vh = v2 / (1 - b2 ** (step + 1))
    w = (w.astype(np.float64) - lr * mh / (np.sqrt(vh) + eps)).astype(np.float32)
    return w, m, v2


def pretrain(rc, w, solvers, steps=200, lr=1e-2, log=print):
    """Fit the scores to the prior by mean squared error."""
    m = np.zeros_like(w, dtype=np.float64)
    v2 = np.zeros_like(w, dtype=np.float64)
    for step in range(steps):
        name, n, s, b, target = solvers[step % len(solvers)]
        sc = s.gnn_scores(w).astype(np.float64)
        ds = 2.0 * (sc - target) / n
        dw = s.gnn_gradient(w, ds.astype(np.float32)).astype(np.float64)
        w, m, v2 = adam_step(w, dw, m, v2, step, lr)
        if step % 20 == 0 or step == steps - 1:
            log(f"[gnn] prior fit step {step} {name}: mse {np.mean((sc - target) ** 2):.4f}")
    return w


def train(steps=300, samples=8, lr=3e-3, seed=0, out=WEIGHTS, small=False, log=print, prior_steps=200):
    rc = _pkg()
    import importlib
    wl = importlib.import_module("rchol_amd.workloads")
    nparam = int(rc.load_library().rc_gnn_weight_count())
    w = init_weights(nparam, seed)
    m = np.zeros_like(w, dtype=np.float64)
    v2 = np.zeros_like(w, dtype=np.float64)
    graphs = training_graphs(wl, small)
    solvers = []
    for name, (n, r, c, v) in graphs:
        s = rc.Solver(device=0, seed=1)
        s.set_triplets(n, r, c, v)
        b = s.make_rhs(rc.RHS_NORMAL_PROJECTED, 42)
        solvers.append((name, n, s, b, prior_scores(n, r, c, v)))
    if prior_steps > 0:
        w = pretrain(rc, w, solvers, prior_steps, log=log)
    rng = np.random.default_rng(seed + 1)
    history = []
    t0 = time.time()
    for step in range(steps):
        name, n, s, b, _ = solvers[step % len(solvers)]
        scores = s.gnn_scores(w).astype(np.float64)
        costs, grads = [], []
        for k in range(samples):
            keys = scores + rng.gumbel(size=n)
            perm = np.argsort(-keys, kind="stable").astype(np.int32)
            s.set_permutation(perm)
            fs = s.factor()
            _, ss = s.solve(b, tol=1e-6)
            costs.append(cost_of(fs, ss))
            grads.append(pl_logp_grad(scores, perm))
        costs = np.array(costs)
        adv = -(costs - costs.mean()) / (costs.std() + 1e-9)
        ds = np.zeros(n)
        for a_k, g_k in zip(adv, grads):
            ds += a_k * g_k
        ds /= samples * n
        # gradient ascent on E[adv * log p]  ==  descent on -that
        dw = -s.gnn_gradient(w, ds.astype(np.float32)).astype(np.float64)
        w, m, v2 = adam_step(w, dw, m, v2, step, lr)
        history.append({"step": step, "graph": name, "cost_mean": float(costs.mean()), "cost_min": float(costs.min())})
        if step % 10 == 0 or step == steps - 1:
            log(f"[gnn] step {step} {name}: cost mean {costs.mean():.4f} min {costs.min():.4f} "
                f"({time.time() - t0:.1f}s)")
    for _, _, s, _, _ in solvers:
        s.close()
    w.tofile(out)
    return w, history


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--samples", type=int, default=8)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--small", action="store_true")
    ap.add_argument("--prior-steps", type=int, default=200)
    ap.add_argument("--eval", default="", help="comma list of eval graphs: grid3d:N or grid2d:N")
    args = ap.parse_args()
    w, hist = train(args.steps, args.samples, args.lr, args.seed, args.out, args.small,
                    log=lambda *a: print(*a, flush=True), prior_steps=args.prior_steps)
    rc = _pkg()
    import importlib
    wl = importlib.import_module("rchol_amd.workloads")
    for spec in [x for x in args.eval.split(",") if x]:
        kind, size = spec.split(":")
        dims = (int(size),) * (3 if kind == "grid3d" else 2)
        res = evaluate(rc, w, wl.grid_laplacian(dims, "loguniform", seed=7))
        print(json.dumps({"eval": spec, **res}), flush=True)


if __name__ == "__main__":
    main()
