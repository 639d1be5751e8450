#!/usr/bin/env python3
"""Write the golden fixtures from the reference's own C code.

Runs only in a container that holds the reference tree: it uses
oracle/_ref/libalbba_ref.so (the reference's biio.h/mmio loader,
convert_csr_to_csc, the SELL-C-sigma builder sequence of sellcs_preprocess
and run_bfs_ref, compiled in place by `make -C oracle ref`).  The fixtures
store inputs and expected outputs only:

  golden_mtx.npz     per .mtx file under mtx/: read_Dmatrix outputs and the
                     transpose convert_csr_to_csc gives
  golden_graphs.npz  per small graph: the CSR working matrix, the reference
                     SELL-C-sigma arrays at several (C, sigma) and the
                     sort-by-first-column variant, and run_bfs_ref's
                     per-iteration frontier counts from vertex 0

Regenerate with:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import albba_amd as A  # noqa: E402
from oracle import oracle as O  # noqa: E402

MTX = ["general_real", "symmetric_pattern", "integer_dups", "hermitian_complex", "skew",
       "mixed_case_comments"]
SELL_CASES = [(256, 16384), (64, 4096), (64, 64), (32, 64), (128, 256), (16, 32)]


def sym_from_edges(n, edges):
    rows = [[] for _ in range(n)]
    for u, v in edges:
        if u != v:
            rows[u].append(v)
            rows[v].append(u)
    rows = [sorted(set(r)) for r in rows]
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    col = np.array([c for r in rows for c in r], np.int32)
    return rp, col


def small_graphs():
    gs = {}
    for name, g in [("rmat8", A.Graph.rmat(8, 8, seed=21, perm_seed=22)),
                    ("rmat10", A.Graph.rmat(10, 16, seed=1, perm_seed=2)),
                    ("grid16", A.Graph.grid(16, 16, 0.1, seed=4, perm_seed=2)),
                    ("er500", A.Graph.er(500, 4.0, seed=7, perm_seed=8))]:
        gs[name] = (g.row_ptr.copy(), g.col.copy())
    gs["path50"] = sym_from_edges(50, [(i, i + 1) for i in range(49)])
    gs["star200"] = sym_from_edges(200, [(0, i) for i in range(1, 200)])
    gs["two_comp"] = sym_from_edges(40, [(i, (i * 7 + 3) % 20) for i in range(20)] +
                                    [(20 + i, 20 + (i * 5 + 1) % 20) for i in range(20)])
    for name in ("general_real", "mixed_case_comments"):
        g = A.Graph.read(os.path.join(HERE, "mtx", name + ".mtx"))
        gs["dir_" + name] = (g.row_ptr.copy(), g.col.copy())
    return gs


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libalbba_ref.so is missing: run `make -C oracle ref` next to the "
                 "reference tree")
    out = {}
    for name in MTX:
        r = O.ref_read(os.path.join(HERE, "mtx", name + ".mtx"))
        nr, nc, nnz, rp, ci, va, sym = r
        out[f"{name}__shape"] = np.array([nr, nc, nnz, sym], np.int64)
        out[f"{name}__row_ptr"] = rp
        out[f"{name}__col_idx"] = ci
        out[f"{name}__val"] = va
        cp, ri = O.ref_csc(nr, nc, rp, ci)
        out[f"{name}__csc_ptr"] = cp
        out[f"{name}__csc_idx"] = ri
    np.savez_compressed(os.path.join(HERE, "golden_mtx.npz"), **out)

    out = {}
    for name, (rp, col) in small_graphs().items():
        out[f"{name}__rp"] = rp
        out[f"{name}__col"] = col
        for C_, sigma in SELL_CASES:
            s = O.ref_sell(rp.astype(np.int32), col, C_, sigma)
            for k, v in s.items():
                out[f"{name}__sell_{C_}_{sigma}__{k}"] = np.asarray(v)
        s = O.ref_sell(rp.astype(np.int32), col, 256, 16384, sort_flag=True)
        for k, v in s.items():
            out[f"{name}__sellsorted_256_16384__{k}"] = np.asarray(v)
        fr, sv = O.ref_bfs_frontiers(rp.astype(np.int32), col)
        out[f"{name}__bfs_ref_frontier"] = fr
        out[f"{name}__bfs_ref_sum_vex"] = np.array([sv], np.int64)
    np.savez_compressed(os.path.join(HERE, "golden_graphs.npz"), **out)
    print("wrote golden_mtx.npz and golden_graphs.npz")


if __name__ == "__main__":
    main()
