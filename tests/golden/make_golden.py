"""Writes the golden fixtures of tests/golden/. Run from the repository root after `make -C oracle ref`.

  tile_policy.json : optimal_tile_size(dim, curr, max) of the reference, computed by the
                     reference's own mm_handle.cpp compiled into oracle/_ref, on the
                     configurations of its tests, CI and benchmarks plus a sweep.
  geometry.json    : tiled_matrix geometry (tile counts, extents, offsets) of the reference
                     (tiled_matrix.cpp compiled into oracle/_ref).
  packing_4x4.json : the 4x4 packing example of Tiled-MM src/Tiled-MM/tiled_mm.hpp:20-35, copied
                     from the text of that comment (host array and the four packed tiles).
  gemm_kat.json    : small integer known-answer GEMMs; expected values by exact integer
                     arithmetic in Python (independent of any BLAS).
"""
import ctypes
import json
import os
import random
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def tile_policy():
    R = oracle.ref()
    cases = [(2048, 0, 512), (40000, 0, 8192), (24000, 0, 6144), (32000, 0, 6144), (50, 0, 4), (200, 0, 4),
             (21, 0, 4), (5, 0, 4), (2, 0, 4), (1000, 0, 5000), (10000, 0, 5000), (1234, 0, 5000),
             (4567, 0, 5000), (1357, 0, 5000), (12345, 0, 5000), (23456, 0, 5000), (67891, 0, 5000),
             (4000, 0, 5000), (32000, 0, 5000), (40000, 5000, 5000), (7, 7, 7), (97, 0, 10), (97, 10, 10)]
    rnd = random.Random(1)
    for _ in range(300):
        mx = rnd.choice([3, 4, 16, 100, 512, 1000, 4096, 5000, 6144, 8192])
        dim = rnd.randint(1, 5 * mx)
        cur = rnd.choice([0, mx, mx // 2, rnd.randint(0, mx)])
        cases.append((dim, cur, mx))
    return [[d, c, m, R.ref_optimal_tile_size(d, c, m)] for d, c, m in cases]


def geometry():
    R = oracle.ref()
    out = []
    rnd = random.Random(2)
    shapes = [(50, 21, 4, 4), (5, 2, 4, 4), (1234, 1357, 500, 300), (4, 4, 2, 2), (7, 3, 10, 10)]
    for _ in range(60):
        r, c = rnd.randint(1, 300), rnd.randint(1, 300)
        shapes.append((r, c, rnd.randint(1, 80), rnd.randint(1, 80)))
    buf = (ctypes.c_longlong * 5)()
    for rows, cols, tr, tc in shapes:
        ld = rows + rnd.randint(0, 5)
        ntr = -(-rows // min(tr, rows))
        ntc = -(-cols // min(tc, cols))
        for ti in sorted({0, ntr - 1, ntr // 2}):
            for tj in sorted({0, ntc - 1, ntc // 2}):
                R.ref_tile_geometry(rows, cols, ld, tr, tc, ti, tj, buf)
                out.append({"rows": rows, "cols": cols, "ld": ld, "tile": [tr, tc], "coord": [ti, tj],
                            "num_tiles": [buf[0], buf[1]], "extent": [buf[2], buf[3]], "offset": buf[4]})
    return out


def packing():
    # tiled_mm.hpp:20-35: matrix rows "1 2 5 4 / 4 2 3 5 / 1 6 7 4 / 9 0 1 2", column-major host
    # array and the packed 2x2 tiles, tile id = <row><col>.
    return {"host": [1, 4, 1, 9, 2, 2, 6, 0, 5, 3, 7, 1, 4, 5, 4, 2], "rows": 4, "cols": 4, "tile": [2, 2],
            "tiles": {"00": [1, 4, 2, 2], "10": [1, 9, 6, 0], "01": [5, 3, 4, 5], "11": [7, 1, 4, 2]}}


def gemm_kat():
    rnd = random.Random(3)
    cases = []
    for (m, n, k, ta, tb, alpha, beta) in [(5, 2, 2, "N", "N", 1, 0), (7, 5, 9, "T", "N", 2, 3), (6, 4, 3, "N", "T", 1, 1),
                                          (4, 6, 5, "T", "T", -1, 2), (3, 3, 1, "N", "N", 1, 1), (9, 7, 11, "N", "N", 1, -2)]:
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        A = [[rnd.randint(0, 9) for _ in range(ac)] for _ in range(ar)]
        B = [[rnd.randint(0, 9) for _ in range(bc)] for _ in range(br)]
        C = [[rnd.randint(0, 9) for _ in range(n)] for _ in range(m)]
        opA = (lambda i, p: A[i][p]) if ta == "N" else (lambda i, p: A[p][i])
        opB = (lambda p, j: B[p][j]) if tb == "N" else (lambda p, j: B[j][p])
        E = [[alpha * sum(opA(i, p) * opB(p, j) for p in range(k)) + beta * C[i][j] for j in range(n)] for i in range(m)]
        cases.append({"m": m, "n": n, "k": k, "ta": ta, "tb": tb, "alpha": alpha, "beta": beta,
                      "A": A, "B": B, "C": C, "expected": E})
    return cases


if __name__ == "__main__":
    with open(os.path.join(OUT, "tile_policy.json"), "w") as f:
        json.dump(tile_policy(), f)
    with open(os.path.join(OUT, "geometry.json"), "w") as f:
        json.dump(geometry(), f)
    with open(os.path.join(OUT, "packing_4x4.json"), "w") as f:
        json.dump(packing(), f, indent=1)
    with open(os.path.join(OUT, "gemm_kat.json"), "w") as f:
        json.dump(gemm_kat(), f)
    print("golden fixtures written to", OUT)
