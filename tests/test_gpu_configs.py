"""GPU parity at the full sizes of the three large workloads.

- RMAT scale 24 (edge factor 16, a/b/c = .57/.19/.19), 64 held-out roots:
  the heuristic and the learned switch, levels and min-id parents bit for
  bit against the oracle's ALBBA restatement (bfs/src/bfs_SELLCSv2.c:
  126-765); the heuristic's per-level direction against the mode the
  restatement records (bfs_SELLCSv2.c:282-290, 302, 344, 356).
- 4096 x 4096 grid with 10% of the edges removed, 32 roots from the largest
  component: against the oracle's queue BFS (thousands of levels).
- The 256-instance batched mix in its autotuned layouts, one launch per
  policy: every instance against the oracle's queue BFS.

The oracle runs on the host's cores (OpenMP inside the ALBBA restatement; a
thread pool over roots or instances for the queue BFS, whose ctypes calls
release the GIL).
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import albba_amd as A
from albba_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu
SELECTOR = os.path.join(os.path.dirname(A.__file__), "selector.txt")


def host_threads() -> int:
    return max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))))


def heuristic_modes(frontier, n):
    """The reference rule on frontier sizes: pull (the slice-bypass pull)
    when F > (int)(0.04 n), else push."""
    thr = int(float(n) * 0.04)
    return np.where(np.asarray(frontier) > thr, A.CHOICE_PULL_SB, A.CHOICE_PUSH)


def test_config4_rmat24_learned_and_heuristic(ctx):
    dg, g = A.DeviceGraph.generate(ctx, "rmat", 64, 4096, 1, **W.RMAT24)
    roots = g.pick_roots(*W.RMAT24_ROOTS)
    assert len(roots) == 64
    sel = A.Selector(ctx)
    sel.load(SELECTOR)
    s = O.build_sell(g.row_ptr, g.col, 256, 16384)
    try:
        for i, root in enumerate(roots):
            if i % 16 == 0:
                print(f"rmat24: root {i}/64", flush=True)
            ref = O.albba_bfs_on(s, g.n, int(root), tasks=host_threads())
            want_modes = np.where(ref["mode"] == 1, A.CHOICE_PULL_SB, A.CHOICE_PUSH)
            for name, pol, sl in (("heuristic", A.SWITCH_HEURISTIC, None),
                                  ("learned", A.SWITCH_LEARNED, sel)):
                o = dg.bfs(int(root), policy=pol, selector=sl)
                np.testing.assert_array_equal(o["level"], ref["level"], err_msg=f"{name} root {root}")
                np.testing.assert_array_equal(o["parent"], ref["parent"], err_msg=f"{name} root {root}")
                np.testing.assert_array_equal(o["frontier"], ref["frontier"])
                if pol == A.SWITCH_HEURISTIC:
                    np.testing.assert_array_equal(o["choice"], want_modes, err_msg=f"root {root}")
            if i < 2:
                assert O.validate(g.row_ptr, g.col, int(root), o["level"], o["parent"]) == 0
    finally:
        O.free_sell(s)
        sel.close()
        dg.close()


def test_config2_grid4096_32_roots(ctx):
    dg, g = A.DeviceGraph.generate(ctx, "grid", 64, 4096, 1, **W.GRID4096)
    rp, col = g.row_ptr, g.col
    roots = g.pick_roots(*W.GRID4096_ROOTS, largest_cc=True)
    assert len(roots) == 32
    workers = host_threads()
    try:
        with ThreadPoolExecutor(workers) as ex:
            for k in range(0, len(roots), workers):
                print(f"grid4096: roots {k}..{k + workers - 1}", flush=True)
                chunk = roots[k:k + workers]
                refs = list(ex.map(lambda r: O.bfs_queue(rp, col, int(r)), chunk))
                for j, (root, (lv, pa)) in enumerate(zip(chunk, refs)):
                    o = dg.bfs(int(root), capacity=1 << 14)
                    np.testing.assert_array_equal(o["level"], lv, err_msg=f"root {root}")
                    np.testing.assert_array_equal(o["parent"], pa, err_msg=f"root {root}")
                    hist = np.bincount(lv[lv >= 0])
                    np.testing.assert_array_equal(o["frontier"], hist)
                    assert o["num_levels"] > 1000
                    np.testing.assert_array_equal(o["choice"], heuristic_modes(hist, g.n))
                    if k + j < 2:
                        assert O.validate(rp, col, int(root), o["level"], o["parent"]) == 0
                del refs
    finally:
        dg.close()


def test_config3_batch256_every_instance(ctx):
    specs = W.batch_specs(256)
    tuned = W.tune_layouts(ctx, specs)
    sel = A.Selector(ctx)
    sel.load(SELECTOR)
    dgs, roots, refs, mapping, sizes = [], [], [], [], []
    workers = host_threads()
    try:
        with ThreadPoolExecutor(workers) as ex:
            for i, (kind, p) in enumerate(specs):
                cfg = W.tuned_config(tuned, kind, p)
                dg, g = W.generate_instance(ctx, kind, p, cfg)
                r = W.root_of(kind, g, i)

                def oracle(g=g, r=r):
                    lv, pa = O.bfs_queue(g.row_ptr, g.col, r)
                    deg = np.diff(g.row_ptr)
                    return lv, pa, int(deg[lv >= 0].sum())

                refs.append(ex.submit(oracle))
                if (i + 1) % 32 == 0:
                    print(f"batch256: {i + 1} instances generated", flush=True)
                dgs.append(dg)
                roots.append(r)
                mapping.append(cfg["mapping"])
                sizes.append((g.n, g.nnz))
            refs = [f.result() for f in refs]
        b = A.Batch(ctx, dgs)
        try:
            for name, pol, sl in (("heuristic", A.SWITCH_HEURISTIC, None),
                                  ("learned", A.SWITCH_LEARNED, sel)):
                out = b.run(np.array(roots, np.int32), policy=pol, selector=sl, mapping=mapping)
                print(f"batch256: {name} launch {out['total_ms']:.1f} ms", flush=True)
                for i, (lv, pa, redges) in enumerate(refs):
                    msg = f"{name}: instance {i} {specs[i]}"
                    assert out["error"][i] == 0, msg
                    glv, gpa = b.results(i)
                    np.testing.assert_array_equal(glv, lv, err_msg=msg)
                    np.testing.assert_array_equal(gpa, pa, err_msg=msg)
                    assert out["reached"][i] == int((lv >= 0).sum()), msg
                    assert out["num_levels"][i] == int(lv.max()) + 1, msg
                    assert out["edges_in_component"][i] == redges, msg
                assert out["total_ms"] > 0
        finally:
            b.close()
    finally:
        for dg in dgs:
            dg.close()
        sel.close()


def test_column_bypass_past_2_24_vertices(ctx):
    """The column-bypass pull's LDS frontier summary (a bit per 64-vertex
    word, 34 KB here) on an ER graph of 2^24 + 2^20 vertices, every level on
    that pull and under the heuristic, against the oracle's queue BFS."""
    n = (1 << 24) + (1 << 20)
    dg, g = A.DeviceGraph.generate(ctx, "er", 64, 4096, 1, n=n, mean_degree=4.0, seed=501, perm_seed=502)
    try:
        root = int(g.pick_roots(1, 503)[0])
        lv, pa = O.bfs_queue(g.row_ptr, g.col, root)
        for pol in (A.SWITCH_FIXED, A.SWITCH_HEURISTIC):
            o = dg.bfs(root, policy=pol, fixed_choice=A.CHOICE_PULL_CB)
            np.testing.assert_array_equal(o["level"], lv)
            np.testing.assert_array_equal(o["parent"], pa)
    finally:
        dg.close()
