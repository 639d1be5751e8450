#!/usr/bin/env python3
"""Benchmark: harmonic-mean GTEPS over 64 roots on RMAT scale 24 (edge factor
16, a=0.57 b=0.19 c=0.19), learned vs heuristic direction switch, levels and
parents bit-exact between the two.

A step is one BFS from each of the 64 roots (one batch).  GTEPS follows the
reference (bfs/src/report_results.c:36): stored directed entries / BFS time.
The graph is resident in HBM before timing starts; the timed region covers
the BFS level loops only.  Prints one JSON line on stdout; progress goes to
stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "harmonic-mean GTEPS, 64 roots, RMAT s24; learned vs heuristic switch, bit-exact"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
SELECTOR = os.path.join(ROOT, "albba_amd", "selector.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pull_bytes(entries, words, n, n_new):
    # 4 B per column entry of the scanned (not bypassed) slices -- every
    # stored entry of the words scanned, the entries a full scan reads --,
    # 8 B slice metadata + 2 x (C/8) B visited read/write per scanned 64-lane
    # word, one frontier-bitmap pass, 8 B of level+parent per discovered
    # vertex.  With `entries` = the entries the early-exit scan gathered, the
    # same formula gives the bytes this kernel cannot avoid.
    return 4 * entries + 8 * words + 16 * words + n / 8 + 8 * n_new


def push_bytes(queue, edges, n, n_new):
    # frontier list + row pointer pairs + adjacency + outputs + visited pass
    return 4 * queue + 8 * queue + 4 * edges + 8 * n_new + n / 8


# kernel (as rocprofv3 names it) that does the main work of each choice
KERNEL_OF_CHOICE = {0: "k_push_expand", 1: "k_pull<64, 0>", 2: "k_pull<64, 1>",
                    3: "k_pull<64, 2>"}


def run_policy(dg, roots, policy, sel, steps, scripts=None):
    """Time `steps` passes over all roots (no per-level events); returns the
    per-root mean BFS ms (events around each BFS), the mean step ms and the
    stored entries of each root's component.  scripts[i]: the per-level
    choices of root i (the scripted policy)."""
    per_root = np.zeros(len(roots))
    comp = np.zeros(len(roots))
    step_ms = []
    for s in range(steps):
        dg.ctx.synchronize()
        t0 = time.perf_counter()
        # one native call runs the roots back to back (albba_bfs_multi)
        o = dg.bfs_many(roots, policy=policy, selector=sel, scripts=scripts)
        dg.ctx.synchronize()
        per_root += o["bfs_ms"] / steps
        comp[:] = o["edges_in_component"]
        step_ms.append((time.perf_counter() - t0) * 1e3)
        log(f"  step {s}: {step_ms[-1]:.1f} ms")
    return per_root, float(np.mean(step_ms)), comp


def account_kernels(dg, roots, policy, sel, n):
    """One pass with HIP events around every level's main kernel (on the
    library's stream): algorithmic bytes and time per kernel."""
    import albba_amd as A

    prof = {}
    for r in roots:
        o = dg.bfs(int(r), policy=policy, selector=sel, time_levels=True, results=False)
        L = o["num_levels"]
        for lv in range(L):
            n_new = o["frontier"][lv + 1] if lv + 1 < L else 0
            ch = int(o["choice"][lv])
            if ch == A.CHOICE_PUSH:
                b = bg = push_bytes(o["words"][lv], o["entries"][lv], n, n_new)
            else:
                # the scanned words' stored entries where the library counted
                # them (list pulls), else the gathered ones
                sw = o["swept"][lv] if o["swept"][lv] >= 0 else o["entries"][lv]
                b = pull_bytes(sw, o["words"][lv], n, n_new)
                bg = pull_bytes(o["entries"][lv], o["words"][lv], n, n_new)
            p = prof.setdefault(KERNEL_OF_CHOICE[ch], {"ms": 0.0, "bytes": 0.0, "bytes_gathered": 0.0,
                                                       "launches": 0})
            p["ms"] += float(o["kernel_ms"][lv])
            p["bytes"] += float(b)
            p["bytes_gathered"] += float(bg)
            p["launches"] += 1
    return prof


def traffic_from_profile(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration)."""
    path = os.path.join(ROOT, "profiles", "k_pull_rmat_s24_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split("|")
        if len(parts) > 3 and parts[1].strip() == kernel:
            try:
                return float(parts[-2])
            except ValueError:
                return None
    return None


def compare_builders(A, ctx, g):
    """The SELL layout (C=64, sigma=4096, pad=1) built on the host (C +
    OpenMP, then uploaded) and on the device from the CSR: times, and whether
    every array is identical."""
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t = time.perf_counter()
    sell = g.sellcs(64, 4096, 1)
    host_ms = (time.perf_counter() - t) * 1e3
    t = time.perf_counter()
    dh = A.DeviceGraph(ctx, sell)
    upload_ms = (time.perf_counter() - t) * 1e3
    ref = sell.arrays()
    del sell
    dh.close()
    t = time.perf_counter()
    dd = A.DeviceGraph.build(ctx, g, 64, 4096, 1)
    wall_ms = (time.perf_counter() - t) * 1e3
    cold = dd.build_report
    dd.close()
    # a second build: the first one also pays for mapping fresh device memory
    dd = A.DeviceGraph.build(ctx, g, 64, 4096, 1)
    rep = dd.build_report
    got = dd.layout().arrays()
    same = all((v is None and got[k] is None) or (v is not None and np.array_equal(v, got[k]))
               for k, v in ref.items())
    # bytes the build must move at least: read the CSR columns, gather their
    # positions, write the columns and the row-major copy
    padded = int(ref["slice_pointers"][-1])
    alg_bytes = 4 * g.nnz * 3 + 4 * padded
    dd.close()
    del ref, got
    out = {"host_build_ms": round(host_ms, 1), "host_threads": threads,
           "host_upload_ms": round(upload_ms, 1),
           "device_csr_upload_ms": round(rep["upload_ms"], 1),
           "device_build_ms": round(rep["build_ms"], 2),
           "device_first_build_ms": round(cold["build_ms"], 2), "device_first_wall_ms": round(wall_ms, 1),
           "device_build_gbps": round(alg_bytes / (rep["build_ms"] * 1e-3) / 1e9, 1),
           "identical": bool(same)}
    log(f"SELL build: host {host_ms:.0f} ms + upload {upload_ms:.0f} ms; device {rep['build_ms']:.1f} ms "
        f"(+ CSR upload {rep['upload_ms']:.0f} ms), identical: {same}")
    return out


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_alba_sample(g, roots, budget_s=12.0, max_roots=4, reps=3, gpu_levels=None):
    """The oracle's C + OpenMP restatement of ALBBA (reference layout C=256,
    sigma=16384) on a bounded sample of roots: the reference's fixed split
    into 8 tasks (bfs/src/bfs_SELLCSv2.c:213-214) and tasks = threads, the
    median of `reps` runs per root, harmonic-mean GTEPS.  gpu_levels[i] =
    (level, parent) of root i on the GPU, compared when given."""
    from oracle import oracle as O
    import albba_amd as A

    threads = host_threads()
    t = time.time()
    s = O.build_sell(g.row_ptr, g.col, 256, 16384)
    log(f"cpu baseline: oracle SELL build {time.time() - t:.1f}s")
    variants = {"tasks_8": 8, "tasks_threads": threads}
    ms = {k: [] for k in variants}
    parity, spent = True, 0.0
    try:
        for i, r in enumerate(roots[:max_roots]):
            for name, tasks in variants.items():
                runs = []
                for _ in range(reps):
                    t0 = time.perf_counter()
                    res = O.albba_bfs_on(s, g.n, int(r), tasks=tasks)
                    runs.append((time.perf_counter() - t0) * 1e3)
                    if gpu_levels is not None and i in gpu_levels:
                        parity &= bool(np.array_equal(res["level"], gpu_levels[i][0]) and
                                       np.array_equal(res["parent"], gpu_levels[i][1]))
                ms[name].append(float(np.median(runs)))
                spent += sum(runs) / 1e3
            log(f"cpu baseline: root {r}: 8 tasks {ms['tasks_8'][-1]:.0f} ms, "
                f"{threads} tasks {ms['tasks_threads'][-1]:.0f} ms (median of {reps})")
            if spent > budget_s:
                break
    finally:
        O.free_sell(s)
    out = {k: {"hm_gteps": round(A.harmonic_mean([A.gteps(g.nnz, x) for x in v]), 4),
               "median_ms_per_root": [round(x, 2) for x in v]} for k, v in ms.items()}
    return out, len(ms["tasks_8"]), parity


def cpu_baseline(g, roots, gpu_levels):
    """CPU baseline of the headline workload (bounded sample of its roots)."""
    threads = host_threads()
    var, k, parity = cpu_alba_sample(g, roots, gpu_levels=gpu_levels)
    return {"value": var["tasks_threads"]["hm_gteps"], "unit": "GTEPS", "cores": threads,
            "kind": "port", "cpu_model": cpu_model(),
            "sample": f"ALBBA level loop restated in C+OpenMP (oracle), the first {k} of the same "
                      f"roots on the same RMAT s24 graph, SELL C=256 sigma=16384, median of 3 runs "
                      f"per root, harmonic mean; value = {threads} tasks (one per thread)",
            "variants": var, "levels_match_gpu": parity}


def batch_results(ctx, count=256, steps=3, which="both"):
    """Batched multi-instance BFS: every instance in one launch.  Layouts from
    the autotuner per (family, size class); the heuristic and the learned
    direction policies on the same instances; sampled instances checked bit
    for bit against the single-BFS path (every instance is checked against
    the oracle by the GPU test of this workload)."""
    import albba_amd as A
    from albba_amd import workloads as W

    specs = W.batch_specs(count)
    sel = None
    if os.path.exists(SELECTOR):
        sel = A.Selector(ctx)
        sel.load(SELECTOR)
    policies = [("heuristic", A.SWITCH_HEURISTIC, None)]
    if sel is not None:
        policies.append(("learned", A.SWITCH_LEARNED, sel))
    if which != "both":
        policies = [p for p in policies if p[0] == which]
    t = time.time()
    tuned = W.tune_layouts(ctx, specs, log=log)
    log(f"autotuning {time.time() - t:.1f}s")
    t = time.time()
    dgs, roots, nnz, kinds, keep = [], [], [], [], {}
    other_roots = [[] for _ in range(steps)]
    for i, (kind, p) in enumerate(specs):
        cfg = W.tuned_config(tuned, kind, p)
        dg, g = W.generate_instance(ctx, kind, p, cfg)
        dgs.append(dg)
        roots.append(W.root_of(kind, g, i))
        for k in range(steps):  # other queries on the same graphs
            other_roots[k].append(W.root_of(kind, g, i + (k + 1) * count))
        nnz.append(g.nnz)
        kinds.append(kind)
        del g
        if i < 3 or i % 64 == 0:
            keep[i] = dg  # checked against the single-BFS path below
        if (i + 1) % 32 == 0:
            log(f"  {i + 1}/{count} instances generated ({time.time() - t:.1f}s)")
    mapping = [W.tuned_config(tuned, k, p)["mapping"] for k, p in specs]
    b = A.Batch(ctx, dgs)
    total = float(np.sum(nnz))
    res, exact = {}, True
    for name, pol, sl in policies:
        # Team sizes come from the previous run's levels and times
        # (albba_batch_run): the first run of a batch, with no history, splits
        # by the instances' shares of the entries; each run on new roots sizes
        # its teams from other queries; the warm-up and the timed steps repeat
        # the same roots.
        b.forget()
        first = b.run(roots, policy=pol, selector=sl, mapping=mapping)
        log(f"  {name} first run (no history): {first['total_ms']:.2f} ms")
        new_ms = []
        for k in range(steps):
            o = b.run(other_roots[k], policy=pol, selector=sl, mapping=mapping)
            new_ms.append(o["total_ms"])
        log(f"  {name} new roots each run: {', '.join(f'{x:.2f}' for x in new_ms)} ms")
        b.run(roots, policy=pol, selector=sl, mapping=mapping)  # warm-up
        runs = []
        for s in range(steps):
            out = b.run(roots, policy=pol, selector=sl, mapping=mapping)
            runs.append(out)
            log(f"  {name} step {s}: {out['total_ms']:.2f} ms")
        for i, dg in keep.items():
            lv, pa = b.results(i)
            o = dg.bfs(roots[i], policy=A.SWITCH_HEURISTIC)
            exact &= bool(np.array_equal(lv, o["level"]) and np.array_equal(pa, o["parent"]))
        ms = float(np.mean([o["total_ms"] for o in runs]))
        done = np.concatenate([o["done_ms"] for o in runs])
        serv = np.concatenate([o["ms"] for o in runs])
        by_kind = {k: round(float(np.mean([runs[-1]["ms"][i] for i in range(count) if kinds[i] == k])), 3)
                   for k in sorted(set(kinds))}
        res[name] = {"gteps": round(total / (ms * 1e6), 3), "ms": round(ms, 3),
                     "first_run_ms": round(first["total_ms"], 3),
                     "new_roots_ms": round(float(np.mean(new_ms)), 3),
                     "latency_ms": {"p50": round(float(np.percentile(done, 50)), 3),
                                    "p99": round(float(np.percentile(done, 99)), 3),
                                    "service_p50": round(float(np.percentile(serv, 50)), 3),
                                    "service_p99": round(float(np.percentile(serv, 99)), 3)},
                     "mean_service_ms_by_kind": by_kind,
                     "levels_max": int(max(o["num_levels"].max() for o in runs)),
                     "pull_levels": int(runs[-1]["pull_levels"].sum()),
                     "push_levels": int(runs[-1]["push_levels"].sum())}
    log(f"sampled instances bit-exact: {exact}")
    b.close()
    for dg in dgs:
        dg.close()
    if sel is not None:
        sel.close()
    return {"instances": count, "rmat": kinds.count("rmat"), "er": kinds.count("er"),
            "grid": kinds.count("grid"), "nnz_stored_total": int(total),
            "layouts": {f"{k[0]}_2^{k[1]}": v for k, v in tuned.items()}, "policies": res,
            "learned_over_heuristic": (round(res["learned"]["gteps"] / res["heuristic"]["gteps"], 4)
                                       if len(res) == 2 else None),
            "sampled_bit_exact": exact}


def run_batch_workload(args):
    import albba_amd as A
    from albba_amd import workloads as W

    ctx = A.Context(0)
    B = batch_results(ctx, args.instances, args.steps, args.policy)
    main_name = "learned" if "learned" in B["policies"] else next(iter(B["policies"]))
    R = B["policies"][main_name]
    line = {
        "metric": f"batched BFS total GTEPS, {B['instances']} instances (RMAT s14-20, ER 2^16-2^20, "
                  "grids 256^2-2048^2), one launch",
        "value": R["gteps"] if B["sampled_bit_exact"] else None, "unit": "GTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": R["ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic: seeded mix (splitmix64 counter RNG), one root per instance",
        "config": {"workload": f"batch{B['instances']}_mix{W.BATCH_MIX_SEED}", "policy": main_name,
                   **{k: B[k] for k in ("instances", "rmat", "er", "grid", "nnz_stored_total", "layouts")}},
        "latency_ms": R["latency_ms"], "policies": B["policies"],
        "learned_over_heuristic": B["learned_over_heuristic"], "sampled_bit_exact": B["sampled_bit_exact"],
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    if not B["sampled_bit_exact"]:
        sys.exit(1)


def run_cpu_workload(args):
    """The CPU baseline of the other configurations (the oracle's C + OpenMP
    restatement of ALBBA, reference layout C=256 sigma=16384, the reference's
    8 tasks and tasks = threads, median of 3 runs per root, harmonic mean),
    each on a bounded sample stated in its entry; plus the dense
    run_bfs_ref-style oracle (bfs/src/bfs_ref.h:26-67) on RMAT s16."""
    from oracle import oracle as O
    import albba_amd as A
    from albba_amd import workloads as W

    threads = host_threads()
    res = {}
    t = time.time()
    # config 0: RMAT s16, ef 16, 16 roots (root seed 3)
    g = A.Graph.rmat(16, 16, 0.57, 0.19, 0.19, seed=1, perm_seed=2)
    roots = g.pick_roots(16, 3)
    var, k, _ = cpu_alba_sample(g, roots, budget_s=30.0, max_roots=16)
    dense = []
    for r in roots[:4]:
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            O.bfs_dense_ref(g.row_ptr, g.col, int(r))
            runs.append((time.perf_counter() - t0) * 1e3)
        dense.append(float(np.median(runs)))
    res["config0_rmat16_16roots"] = {
        "nnz": g.nnz, "roots_timed": k, **var,
        "dense_run_bfs_ref": {"hm_gteps": round(A.harmonic_mean([A.gteps(g.nnz, x) for x in dense]), 4),
                              "median_ms_per_root": [round(x, 2) for x in dense], "threads": 1,
                              "roots_timed": len(dense)}}
    log(f"config 0 done ({time.time() - t:.0f}s): {res['config0_rmat16_16roots']}")
    # config 1: RMAT s22, ef 16, 64 roots (root seed 3): the first 8
    t = time.time()
    g = A.Graph.rmat(22, 16, 0.57, 0.19, 0.19, seed=1, perm_seed=2)
    roots = g.pick_roots(64, 3)
    var, k, _ = cpu_alba_sample(g, roots, budget_s=30.0, max_roots=8)
    res["config1_rmat22_64roots"] = {"nnz": g.nnz, "roots_timed": k, **var}
    log(f"config 1 done ({time.time() - t:.0f}s): {res['config1_rmat22_64roots']}")
    del g
    # config 2: the 4096^2 grid has thousands of levels, each a scan of the
    # layout in the restatement: the first `cap` levels of 2 roots are timed
    # and the per-level time scaled to the root's level count
    t = time.time()
    g = A.Graph.grid(**W.GRID4096)
    roots = g.pick_roots(*W.GRID4096_ROOTS, largest_cc=True)
    s = O.build_sell(g.row_ptr, g.col, 256, 16384)
    cap = 30
    grid = {}
    try:
        levels = [int(O.bfs_queue(g.row_ptr, g.col, int(r))[0].max()) + 1 for r in roots[:2]]
        for name, tasks in (("tasks_8", 8), ("tasks_threads", threads)):
            per_level, est = [], []
            for r, nl in zip(roots[:2], levels):
                runs = []
                for _ in range(3):
                    t0 = time.perf_counter()
                    got = O.albba_bfs_on(s, g.n, int(r), tasks=tasks, capacity=cap + 1, max_levels=cap)
                    runs.append((time.perf_counter() - t0) * 1e3 / got["num_levels"])
                per_level.append(float(np.median(runs)))
                est.append(per_level[-1] * nl)
            grid[name] = {"hm_gteps": round(A.harmonic_mean([A.gteps(g.nnz, x) for x in est]), 6),
                          "ms_per_level": [round(x, 3) for x in per_level],
                          "est_ms_per_root": [round(x, 1) for x in est]}
    finally:
        O.free_sell(s)
    res["config2_grid4096_32roots"] = {"nnz": g.nnz, "roots_timed": 2, "levels": levels,
                                       "sample": f"first {cap} levels of 2 roots, per-level median of 3 "
                                                 "scaled to the root's level count", **grid}
    log(f"config 2 done ({time.time() - t:.0f}s): {res['config2_grid4096_32roots']}")
    del g
    # config 3: the batch mix run instance after instance: 6 RMAT and 6 ER
    # instances in full, 3 grids level-capped as above
    t = time.time()
    lat, nnz_s = [], []
    n_done = {"rmat": 0, "er": 0, "grid": 0}
    for i, (kind, p) in enumerate(W.batch_specs(256)):
        if n_done[kind] >= (3 if kind == "grid" else 6):
            continue
        g = W.make_graph(kind, p)
        r = W.root_of(kind, g, i)
        s = O.build_sell(g.row_ptr, g.col, 256, 16384)
        try:
            if kind == "grid":
                nl = int(O.bfs_queue(g.row_ptr, g.col, r)[0].max()) + 1
                t0 = time.perf_counter()
                got = O.albba_bfs_on(s, g.n, r, tasks=threads, capacity=cap + 1, max_levels=cap)
                ms = (time.perf_counter() - t0) * 1e3 / got["num_levels"] * nl
            else:
                t0 = time.perf_counter()
                O.albba_bfs_on(s, g.n, r, tasks=threads)
                ms = (time.perf_counter() - t0) * 1e3
        finally:
            O.free_sell(s)
        lat.append(ms)
        nnz_s.append(g.nnz)
        n_done[kind] += 1
        if sum(n_done.values()) == 15:
            break
    res["config3_batch256"] = {"instances_timed": len(lat), "by_kind": n_done,
                               "total_gteps_sample": round(float(np.sum(nnz_s)) / (float(np.sum(lat)) * 1e6), 4),
                               "latency_ms_p50": round(float(np.percentile(lat, 50)), 2),
                               "latency_ms_p99": round(float(np.percentile(lat, 99)), 2),
                               "sample": f"{len(lat)} instances one after the other ({threads} tasks): "
                                         "6 RMAT and 6 ER in full, 3 grids level-capped and scaled"}
    log(f"config 3 done ({time.time() - t:.0f}s): {res['config3_batch256']}")
    line = {"metric": "CPU baseline of configs 0-3 (oracle's ALBBA restatement, C + OpenMP)",
            "value": res["config1_rmat22_64roots"]["tasks_threads"]["hm_gteps"], "unit": "GTEPS",
            "n_gpus": 0, "steps": 1, "warmup": 0, "ms_per_step": None, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic, seeded",
            "config": {"workload": "cpu_baseline_configs0-3", "threads": threads, "cpu_model": cpu_model()},
            "cpu": res}
    print(json.dumps(line), flush=True)


def grid_results(ctx, steps=1):
    """High-diameter case: 4-neighbour grid of 4096 x 4096 vertices with 10%
    of the edges removed (seed 4, ids permuted with seed 2), 32 roots from
    the largest component (seed 5), heuristic switch, generated on the
    device; the level loop with the persistent small-frontier kernel, and
    (steps > 1) with one launch sequence per level for comparison."""
    import albba_amd as A
    from albba_amd import workloads as W

    t = time.time()
    dg, g = A.DeviceGraph.generate(ctx, "grid", 64, 4096, 1, **W.GRID4096)
    roots = g.pick_roots(*W.GRID4096_ROOTS, largest_cc=True)
    log(f"grid 4096^2: n={g.n} nnz={g.nnz} ({time.time() - t:.1f}s)")
    res = {}
    for name, cap in (("small_frontier_kernel", None), ("per_level_launches", 0)):
        dg.bfs(int(roots[0]), small_frontier_edges=cap, results=False, capacity=1 << 15)  # warm-up
        ms, levels = [], []
        for r in roots:
            o = dg.bfs(int(r), small_frontier_edges=cap, results=False, capacity=1 << 15)
            ms.append(o["bfs_ms"])
            levels.append(o["num_levels"])
        res[name] = {"hm_gteps": A.harmonic_mean([A.gteps(g.nnz, x) for x in ms]),
                     "mean_ms": float(np.mean(ms)), "levels_mean": float(np.mean(levels)),
                     "us_per_level": float(np.sum(ms) * 1e3 / np.sum(levels))}
        log(f"  {name}: {res[name]}")
        if steps <= 1:
            break
    out = {"n": g.n, "nnz_stored": g.nnz, "roots": len(roots),
           "hm_gteps": round(res["small_frontier_kernel"]["hm_gteps"], 4),
           "detail": {k: {kk: round(vv, 4) for kk, vv in v.items()} for k, v in res.items()}}
    dg.close()
    return out


def rmat_config_results(ctx, scale, nroots, root_seed=3, steps=3):
    """A plain RMAT configuration on the GPU (ef 16, a/b/c = .57/.19/.19,
    seeds 1/2, generated on the device, C = 64, sigma = 4096): harmonic-mean
    GTEPS of the heuristic and the learned switch over the roots (root seed
    3, as the parity tests), event-timed per BFS and wall clock per step of
    one albba_bfs_multi call, and whether both policies give the same levels
    and parents on the first roots."""
    import albba_amd as A

    dg, g = A.DeviceGraph.generate(ctx, "rmat", 64, 4096, 1, scale=scale, edgefactor=16, a=0.57, b=0.19,
                                   c=0.19, seed=1, perm_seed=2)
    roots = g.pick_roots(nroots, root_seed)
    sel = None
    if os.path.exists(SELECTOR):
        sel = A.Selector(ctx)
        sel.load(SELECTOR)
    pols = [("heuristic", A.SWITCH_HEURISTIC, None)] + ([("learned", A.SWITCH_LEARNED, sel)] if sel else [])
    out = {"n": g.n, "nnz_stored": g.nnz, "roots": len(roots)}
    exact = True
    for r in roots[:4]:
        ref = dg.bfs(int(r))
        for _, pol, sl in pols[1:]:
            o = dg.bfs(int(r), policy=pol, selector=sl)
            exact &= bool(np.array_equal(o["level"], ref["level"]) and np.array_equal(o["parent"], ref["parent"]))
    for name, pol, sl in pols:
        per_root, step_ms, _ = run_policy(dg, roots, pol, sl, steps)
        out[name] = {"hm_gteps_events": round(A.harmonic_mean([A.gteps(g.nnz, x) for x in per_root]), 3),
                     "wall_gteps": round(len(roots) * g.nnz / (step_ms * 1e6), 3)}
    out["bit_exact"] = exact
    if sel is not None:
        sel.close()
    dg.close()
    return out


def run_grid_workload(args):
    import albba_amd as A

    ctx = A.Context(0)
    R = grid_results(ctx, args.steps)
    line = {
        "metric": "harmonic-mean GTEPS, 32 roots, 4096^2 grid (10% edges removed), heuristic switch",
        "value": R["hm_gteps"], "unit": "GTEPS", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(R["detail"]["small_frontier_kernel"]["mean_ms"] * R["roots"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic: seeded grid, permuted ids",
        "config": {"workload": "grid4096_p0.1_32roots", "n": R["n"], "nnz_stored": R["nnz_stored"],
                   "roots": R["roots"], "layout": "SELL-C-sigma C=64 sigma=4096 pad=1"},
        "detail": R["detail"],
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def run_load_workload(args):
    """.cbd file -> device graph ready for BFS (SELL C=64, sigma=4096,
    pad=1): the host path (read_Dmatrix + the W = A^T transpose + the host
    SELL build + upload) against albba_graph_load (file streamed into device
    memory, transpose and build on the device).  The file is written first,
    so both read it from the page cache.  RMAT of --scale (default 22 here)."""
    import shutil
    import tempfile

    import albba_amd as A

    scale = args.scale if args.scale != 24 else 22
    g = A.Graph.rmat(scale, 16, 0.57, 0.19, 0.19, seed=1, perm_seed=2)
    d = tempfile.mkdtemp(prefix="albba_load_")
    path = os.path.join(d, f"rmat{scale}.cbd")
    try:
        g.write_cbd(path)
        size = os.path.getsize(path)
        log(f"wrote {path}: {size / 2**30:.2f} GiB")
        ctx = A.Context(0)
        runs = {"host": [], "device": []}
        same = None
        for rep_i in range(max(args.steps, 1) + 1):  # the first pair is a warm-up
            t = time.perf_counter()
            h = A.Graph.read(path)
            t_read = time.perf_counter()
            sell = h.sellcs(64, 4096, 1)
            t_build = time.perf_counter()
            dh = A.DeviceGraph(ctx, sell)
            t_up = time.perf_counter()
            host = {"read_ms": (t_read - t) * 1e3, "build_ms": (t_build - t_read) * 1e3,
                    "upload_ms": (t_up - t_build) * 1e3, "total_ms": (t_up - t) * 1e3}
            t = time.perf_counter()
            dd = A.DeviceGraph.load(ctx, path, 64, 4096, 1)
            dev = dict(dd.build_report, wall_ms=(time.perf_counter() - t) * 1e3)
            if same is None:
                ref, got = sell.arrays(), dd.layout().arrays()
                same = all((v is None and got[k] is None) or
                           (v is not None and np.array_equal(v, got[k])) for k, v in ref.items())
                del ref, got
            del sell, h
            dh.close()
            dd.close()
            log(f"  host {host['total_ms']:.0f} ms, device {dev['wall_ms']:.0f} ms ({dev})")
            if rep_i > 0:
                runs["host"].append(host)
                runs["device"].append(dev)
        mean = {k: {kk: round(float(np.mean([r[kk] for r in v])), 2) for kk in v[0]}
                for k, v in runs.items()}
        line = {
            "metric": f"ms from a .cbd file to a device graph ready for BFS, RMAT s{scale}",
            "value": mean["device"]["wall_ms"], "unit": "ms", "n_gpus": 1, "steps": len(runs["host"]),
            "warmup": 1, "ms_per_step": mean["device"]["wall_ms"], "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "int32",
            "data": "synthetic: seeded RMAT written as .cbd (page cache)",
            "config": {"workload": f"cbd_load_rmat_s{scale}", "n": g.n, "nnz_stored": g.nnz,
                       "file_bytes": size, "layout": "SELL-C-sigma C=64 sigma=4096 pad=1"},
            "host_path_ms": mean["host"]["total_ms"], "detail": mean, "identical": bool(same),
        }
        print(json.dumps(line), flush=True)
        ctx.close()
    finally:
        shutil.rmtree(d, ignore_errors=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--roots", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-other-configs", action="store_true",
                    help="headline only: skip the grid and batch measurements")
    ap.add_argument("--workload", choices=["rmat24", "grid", "batch", "load", "cpu"], default="rmat24",
                    help="rmat24: the headline metric; grid: the 4096^2 high-diameter case; "
                         "batch: the batched multi-instance BFS; load: .cbd file to device graph; "
                         "cpu: the CPU baseline of configs 0-3")
    ap.add_argument("--instances", type=int, default=256, help="batch workload size")
    ap.add_argument("--policy", choices=["both", "heuristic", "learned"], default="both",
                    help="run one policy only (profiling); 'both' checks bit-exactness")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU; --gpus {args.gpus} is not supported "
              "(a single BFS would shard as a 1D partition of the SELL slices)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        run_batch_workload(args)
        return
    if args.workload == "grid":
        run_grid_workload(args)
        return
    if args.workload == "load":
        run_load_workload(args)
        return
    if args.workload == "cpu":
        run_cpu_workload(args)
        return

    import albba_amd as A

    ctx = A.Context(0)
    # the graph generated on the device (the host generator's CSR, bit for
    # bit), with a host copy for the roots, the builders' comparison and the
    # CPU baseline
    dg, g = A.DeviceGraph.generate(ctx, "rmat", 64, 4096, 1, scale=args.scale, edgefactor=16,
                                   a=0.57, b=0.19, c=0.19, seed=1, perm_seed=2)
    gen_report = dg.build_report
    log(f"graph RMAT s{args.scale}: n={g.n} nnz={g.nnz}, generated and laid out on the device "
        f"({gen_report}, {dg.device_bytes / 2**30:.2f} GiB on device)")
    roots = g.pick_roots(args.roots, 7)
    sell_build = compare_builders(A, ctx, g)

    sel = None
    if os.path.exists(SELECTOR):
        sel = A.Selector(ctx)
        sel.load(SELECTOR)
    policies = [("heuristic", A.SWITCH_HEURISTIC, None)]
    if sel is not None:
        policies.append(("learned", A.SWITCH_LEARNED, sel))
    if args.policy != "both":
        policies = [p for p in policies if p[0] == args.policy]
        if not policies:
            log(f"policy {args.policy} unavailable (no selector file)")
            sys.exit(1)

    # warmup + bit-exactness: every policy must give identical level/parent
    bit_exact = True
    keep = {}
    for w in range(max(args.warmup, 1)):
        for i, r in enumerate(roots):
            ref = None
            for name, pol, s in policies:
                o = dg.bfs(int(r), policy=pol, selector=s, results=(w == 0))
                if w == 0:
                    if ref is None:
                        ref = (o["level"], o["parent"])
                        if i < 8:
                            keep[i] = ref
                    else:
                        bit_exact &= bool(np.array_equal(ref[0], o["level"]) and
                                          np.array_equal(ref[1], o["parent"]))
    log(f"warmup done, policies bit-exact: {bit_exact}")

    # the per-level oracle choice: every choice timed on a snapshot of each
    # level (the BEST policy), then those choices replayed as a script
    scripts = None
    if args.policy == "both":
        t = time.time()
        scripts = []
        for i, r in enumerate(roots):
            o = dg.bfs(int(r), policy=A.SWITCH_BEST, timing_reps=3, results=(i in keep))
            scripts.append(o["choice"].astype(np.int32))
            if i in keep:
                bit_exact &= bool(np.array_equal(keep[i][0], o["level"]) and
                                  np.array_equal(keep[i][1], o["parent"]))
        for i, r in enumerate(roots):
            o = dg.bfs(int(r), policy=A.SWITCH_SCRIPT, script=scripts[i], results=(i in keep))
            if i in keep:
                bit_exact &= bool(np.array_equal(keep[i][0], o["level"]) and
                                  np.array_equal(keep[i][1], o["parent"]))
        log(f"per-level oracle choices found ({time.time() - t:.1f}s), bit-exact: {bit_exact}")
        policies.append(("oracle_choice", A.SWITCH_SCRIPT, None))

    results = {}
    for name, pol, s in policies:
        log(f"timing policy {name}")
        per_root, step_ms, comp = run_policy(dg, roots, pol, s, args.steps,
                                             scripts if pol == A.SWITCH_SCRIPT else None)
        gt = [A.gteps(g.nnz, ms) for ms in per_root]
        # Graph500 convention: undirected edges of the root's component (each
        # stored twice, no self loops) over the BFS time
        g500 = [A.gteps(c / 2, ms) for c, ms in zip(comp, per_root)]
        results[name] = {"hm_gteps_events": A.harmonic_mean(gt), "step_ms": step_ms,
                         "hm_gteps_graph500": A.harmonic_mean(g500),
                         "wall_gteps": len(roots) * g.nnz / (step_ms * 1e6),
                         "mean_bfs_ms": float(np.mean(per_root))}
        log(f"  {name}: harmonic-mean GTEPS (events) {results[name]['hm_gteps_events']:.2f}, "
            f"wall {results[name]['wall_gteps']:.2f}, step {step_ms:.1f} ms")

    main_name = "learned" if "learned" in results else policies[0][0]
    R = results[main_name]
    pol_main, sel_main = [(p, s) for nm, p, s in policies if nm == main_name][0]
    prof = account_kernels(dg, roots, pol_main, sel_main, g.n)
    dom = max(prof, key=lambda k: prof[k]["ms"])
    P = prof[dom]
    achieved = P["bytes"] / (P["ms"] * 1e-3) if P["ms"] > 0 else 0.0
    copy_bytes = 1 << 30
    copy_ms = ctx.bench_copy(copy_bytes, 20)
    # PCIe-inclusive rate (reported beside, never `value`): BFS plus the copy
    # of level[] and parent[] to host memory through albba_bfs_results
    t0 = time.perf_counter()
    for r in roots[:8]:
        dg.bfs(int(r), policy=pol_main, selector=sel_main, results=True)
    pcie_gteps = min(8, len(roots)) * g.nnz / ((time.perf_counter() - t0) * 1e3 * 1e6)
    achieved_g = P["bytes_gathered"] / (P["ms"] * 1e-3) if P["ms"] > 0 else 0.0
    roofline = {"bound": "hbm", "kernel": dom,
                "achieved": round(achieved / 1e9, 2), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic_from_profile(dom),
                # the same formula over the entries the early-exit scan gathered
                "achieved_gathered": round(achieved_g / 1e9, 2),
                "frac_gathered": round(achieved_g / HBM_PEAK, 4),
                "avg_gathered_bytes_per_launch": round(P["bytes_gathered"] / max(P["launches"], 1)),
                "avg_launch_ms": round(P["ms"] / max(P["launches"], 1), 5),
                "avg_bytes_per_launch": round(P["bytes"] / max(P["launches"], 1)),
                "launches": P["launches"],
                "copy_kernel_gbps": round(2 * copy_bytes / (copy_ms * 1e-3) / 1e9, 1),
                "kernels": {k: {"launches": v["launches"],
                                "avg_ms": round(v["ms"] / max(v["launches"], 1), 5),
                                "gbps": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1)
                                if v["ms"] > 0 else None}
                            for k, v in sorted(prof.items())}}

    def ratio(a, b):
        return round(results[a]["wall_gteps"] / results[b]["wall_gteps"], 4) \
            if a in results and b in results else None

    line = {
        "metric": METRIC, "value": round(R["wall_gteps"], 3), "unit": "GTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(R["step_ms"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic: seeded RMAT (splitmix64 counter RNG, Fisher-Yates permuted ids, "
                "symmetrized, deduplicated, no self loops)",
        "config": {"workload": f"rmat_s{args.scale}_ef16_{len(roots)}roots",
                   "scale": args.scale, "edgefactor": 16, "a": 0.57, "b": 0.19, "c": 0.19,
                   "n": g.n, "nnz_stored": g.nnz, "roots": len(roots), "root_seed": 7,
                   "layout": "SELL-C-sigma C=64 sigma=4096 pad=1", "policy": main_name},
        "heuristic_gteps": (round(results["heuristic"]["wall_gteps"], 3)
                            if "heuristic" in results else None),
        "oracle_choice_gteps": (round(results["oracle_choice"]["wall_gteps"], 3)
                                if "oracle_choice" in results else None),
        "learned_over_heuristic": ratio("learned", "heuristic"),
        "learned_over_oracle_choice": ratio("learned", "oracle_choice"),
        "hm_gteps_events": {k: round(v["hm_gteps_events"], 3) for k, v in results.items()},
        "hm_gteps_graph500": {k: round(v["hm_gteps_graph500"], 3) for k, v in results.items()},
        "bit_exact": bit_exact if len(policies) > 1 else None,
        "pcie_inclusive_gteps": round(pcie_gteps, 3),
        "sell_build": sell_build,
        "graph_generate_ms": {k: round(v, 1) for k, v in gen_report.items()},
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(g, roots, keep)
    if sel is not None:
        sel.close()
    dg.close()
    del g
    if not args.no_other_configs:
        # the other configurations, measured in the same run: RMAT s16 and s22
        # on the GPU, the high-diameter grid and the batched mix
        line["config0_rmat16_gpu"] = rmat_config_results(ctx, 16, 16)
        log(f"config 0 on the GPU: {line['config0_rmat16_gpu']}")
        line["config1_rmat22_gpu"] = rmat_config_results(ctx, 22, 64)
        log(f"config 1 on the GPU: {line['config1_rmat22_gpu']}")
        line["config2_grid4096"] = grid_results(ctx)
        line["config3_batch256"] = batch_results(ctx, 256, 2)
    # a result that is not bit-exact carries no value and fails the run
    ok = (bit_exact is not False and line.get("cpu_baseline", {}).get("levels_match_gpu", True) and
          line.get("config3_batch256", {}).get("sampled_bit_exact", True) and
          line.get("config0_rmat16_gpu", {}).get("bit_exact", True) and
          line.get("config1_rmat22_gpu", {}).get("bit_exact", True))
    if not ok:
        line["value"] = None
        line["error"] = "levels or parents differ between policies or from the oracle"
    print(json.dumps(line), flush=True)
    ctx.close()
    if not ok:
        sys.exit(1)


if __name__ == "__main__":
    main()
