"""CPU oracle of the ordering network (test infrastructure only).

Float64 forward and backward of the network in
parallel-randomized-cholesky_amd/csrc/device/gnn.hip, with the same bfloat16
rounding of weights and layer inputs (round to nearest even) when
`bf16=True`; with `bf16=False` it is the exact float64 network, whose backward
pass is pinned by finite differences in tests/test_gnn.py.

Weight layout: per layer Ws (64 x in), Wn (64 x in), b (64) for in = 2, 64, 64;
then wo (64), bo (1).
"""
import numpy as np
import scipy.sparse as sp

H = 64
LAYOUT = []
_off = 0
for _din in (2, 64, 64):
    LAYOUT.append((_off, _off + H * _din, _off + 2 * H * _din, _din))
    _off += 2 * H * _din + H
WO = _off
BO = _off + H
NPARAM = BO + 1


def bf16(x):
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def graph(n, rows, cols, vals):
    """Off-diagonal weighted graph of a Laplacian/adjacency given as triplets:
    (A row-normalised, centred inputs x0)."""
    off = rows != cols
    r, c, w = rows[off], cols[off], np.abs(vals[off])
    W = sp.csr_matrix((w, (r, c)), shape=(n, n))
    W.sum_duplicates()
    deg = np.diff(W.indptr)
    wdeg = np.asarray(W.sum(axis=1)).ravel()
    A = sp.diags(np.where(wdeg > 0, 1.0 / np.where(wdeg > 0, wdeg, 1.0), 0.0)) @ W
    x0 = np.zeros((n, 2))
    m = deg > 0
    l0 = np.log(deg[m].astype(np.float64))
    l1 = np.log(wdeg[m])
    x0[m, 0] = l0 - l0.mean()
    x0[m, 1] = l1 - l1.mean()
    return sp.csr_matrix(A), x0.astype(np.float32).astype(np.float64)


def _split(w, layer):
    o, on, ob, din = LAYOUT[layer]
    return w[o:o + H * din].reshape(H, din), w[on:on + H * din].reshape(H, din), w[ob:ob + H]


def forward(A, x0, w, bf16_round=True):
    r = bf16 if bf16_round else (lambda v: np.asarray(v, dtype=np.float64))
    w = np.asarray(w, dtype=np.float32).astype(np.float64)
    cache = {"H": [x0], "M": [], "Z": []}
    Hp = x0
    for layer in range(3):
        Ws, Wn, b = _split(w, layer)
        M = A @ Hp
        Z = r(Hp) @ r(Ws).T + r(M) @ r(Wn).T + b
        Hn = np.maximum(Z, 0.0)
        cache["M"].append(M)
        cache["Z"].append(Z)
        cache["H"].append(Hn)
        Hp = Hn
    s = r(Hp) @ r(w[WO:WO + H]) + w[BO]
    return s, cache


def backward(A, w, cache, ds, bf16_round=True):
    r = bf16 if bf16_round else (lambda v: np.asarray(v, dtype=np.float64))
    w = np.asarray(w, dtype=np.float32).astype(np.float64)
    dw = np.zeros(NPARAM)
    H3 = cache["H"][3]
    dw[WO:WO + H] = r(H3).T @ ds
    dw[BO] = ds.sum()
    dH = np.outer(ds, r(w[WO:WO + H]))
    for layer in (2, 1, 0):
        Ws, Wn, b = _split(w, layer)
        o, on, ob, din = LAYOUT[layer]
        dZ = dH * (cache["Z"][layer] > 0)
        Hp, M = cache["H"][layer], cache["M"][layer]
        dw[o:o + H * din] = (dZ.T @ r(Hp)).ravel()
        dw[on:on + H * din] = (dZ.T @ r(M)).ravel()
        dw[ob:ob + H] = dZ.sum(axis=0)
        if layer > 0:
            dH = dZ @ r(Ws) + A.T @ (dZ @ r(Wn))
    return dw


def init_weights(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    w = np.zeros(NPARAM, dtype=np.float32)
    for layer in range(3):
        o, on, ob, din = LAYOUT[layer]
        std = scale * np.sqrt(2.0 / (2 * din))
        w[o:o + H * din] = rng.normal(0, std, H * din)
        w[on:on + H * din] = rng.normal(0, std, H * din)
        w[ob:ob + H] = 0.01
    w[WO:WO + H] = rng.normal(0, scale * np.sqrt(1.0 / H), H)
    w[BO] = 0.0
    return w
