"""Writes the tiny Matrix Market fixtures under tests/golden/ and their
expected Laplacians / ground-appended matrices, computed from first
principles with numpy (L = D - W for a weighted graph; the SDDM ground append
of write_graph.jl:367-383: appended row/column = -column sums clipped at 0,
corner = their negated sum).  The reference ships no fixtures; these pin the
input path.  Run: python tests/golden/make_golden.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def write_mtx(name, n, entries, field="real", sym="general"):
    with open(os.path.join(HERE, name), "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {field} {sym}\n% fixture\n")
        f.write(f"{n} {n} {len(entries)}\n")
        for r, c, v in entries:
            if field == "pattern":
                f.write(f"{r + 1} {c + 1}\n")
            else:
                f.write(f"{r + 1} {c + 1} {float(v)!r}\n")


def lap_from_edges(n, edges):
    w = np.zeros((n, n))
    for u, v, x in edges:
        w[u, v] += x
        w[v, u] += x
    return np.diag(w.sum(axis=0)) - w


def main():
    exp = {}
    # path 0-1-2-3, general, both triangles, no diagonal
    e = [(0, 1, 1.0), (1, 2, 2.0), (2, 3, 3.0)]
    write_mtx("path4.mtx", 4, [(u, v, -x) for u, v, x in e] + [(v, u, -x) for u, v, x in e])
    exp["path4.mtx"] = lap_from_edges(4, e).tolist()
    # star, centre 0, symmetric header (lower triangle only), positive weights
    e = [(0, k, float(k)) for k in range(1, 5)]
    write_mtx("star5.mtx", 5, [(v, u, x) for u, v, x in e], sym="symmetric")
    exp["star5.mtx"] = lap_from_edges(5, e).tolist()
    # triangle, pattern -> unit weights
    e = [(0, 1, 1.0), (0, 2, 1.0), (1, 2, 1.0)]
    write_mtx("triangle.mtx", 3, [(u, v, 1.0) for u, v, x in e] + [(v, u, 1.0) for u, v, x in e], field="pattern")
    exp["triangle.mtx"] = lap_from_edges(3, e).tolist()
    # 3x3 grid Laplacian with an explicit diagonal, symmetric header
    e = []
    for i in range(3):
        for j in range(3):
            v = 3 * i + j
            if j < 2:
                e.append((v, v + 1, 1.0 + 0.25 * v))
            if i < 2:
                e.append((v, v + 3, 2.0 - 0.125 * v))
    L = lap_from_edges(9, e)
    ent = [(r, c, L[r, c]) for c in range(9) for r in range(c, 9) if L[r, c] != 0.0]
    write_mtx("grid3x3.mtx", 9, ent, sym="symmetric")
    exp["grid3x3.mtx"] = L.tolist()
    # SDDM: path 0-1-2 with diagonal = degree + (0.5, 0, 0.25): ground append
    A = lap_from_edges(3, [(0, 1, 1.0), (1, 2, 1.0)]) + np.diag([0.5, 0.0, 0.25])
    write_mtx("sddm3.mtx", 3, [(r, c, A[r, c]) for c in range(3) for r in range(3) if A[r, c] != 0.0])
    cs = A.sum(axis=0)
    app = np.minimum(-cs, 0.0)
    G = np.zeros((4, 4))
    G[:3, :3] = A
    G[3, :3] = app
    G[:3, 3] = app
    G[3, 3] = -app.sum()
    exp["sddm3.mtx"] = A.tolist()
    exp["sddm3.mtx:grounded"] = G.tolist()
    # not diagonally dominant: column sums negative
    A = lap_from_edges(3, [(0, 1, 1.0), (1, 2, 1.0)]) - np.diag([0.5, 0.5, 0.5])
    write_mtx("nondominant.mtx", 3, [(r, c, A[r, c]) for c in range(3) for r in range(3) if A[r, c] != 0.0])
    # two components: edge 0-1 and triangle 2-3-4
    e = [(0, 1, 1.0), (2, 3, 1.0), (3, 4, 2.0), (2, 4, 3.0)]
    write_mtx("two_components.mtx", 5, [(u, v, -x) for u, v, x in e] + [(v, u, -x) for u, v, x in e])
    exp["two_components.mtx"] = lap_from_edges(5, e).tolist()
    # malformed inputs
    with open(os.path.join(HERE, "nonsquare.mtx"), "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n3 4 1\n1 2 1.0\n")
    with open(os.path.join(HERE, "truncated.mtx"), "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n3 3 4\n1 2 -1.0\n2 1 -1.0\n")
    with open(os.path.join(HERE, "badheader.mtx"), "w") as f:
        f.write("%%MatrixMarket matrix array real general\n2 2\n1\n0\n0\n1\n")
    with open(os.path.join(HERE, "expected.json"), "w") as f:
        json.dump(exp, f, indent=1)


if __name__ == "__main__":
    main()
