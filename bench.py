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
    # 4 B per gathered column entry, 8 B slice metadata + 2 x (C/8) B visited
    # read/write per scanned 64-lane word, one frontier-bitmap pass, 8 B of
    # level+parent per discovered vertex
    return 4 * entries + 8 * words + 16 * words + n / 8 + 8 * n_new


def push_bytes(queue, edges, n, n_new):
    # frontier list + row pointer pairs + adjacency + outputs + visited pass
    return 4 * queue + 8 * queue + 4 * edges + 8 * n_new + n / 8


# kernel (as rocprofv3 names it) that does the main work of each choice
KERNEL_OF_CHOICE = {0: "k_push_expand", 1: "k_pull<64, 0>", 2: "k_pull<64, 1>",
                    3: "k_pull<64, 2>"}


def run_policy(dg, roots, policy, sel, steps):
    """Time `steps` passes over all roots (no per-level events); returns the
    per-root mean BFS ms (events around each BFS), the mean step ms and the
    stored entries of each root's component."""
    per_root = np.zeros(len(roots))
    comp = np.zeros(len(roots))
    step_ms = []
    for s in range(steps):
        dg.ctx.synchronize()
        t0 = time.perf_counter()
        for i, r in enumerate(roots):
            o = dg.bfs(int(r), policy=policy, selector=sel, results=False)
            per_root[i] += o["bfs_ms"] / steps
            comp[i] = o["edges_in_component"]
        dg.ctx.synchronize()
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
            b = (push_bytes(o["words"][lv], o["entries"][lv], n, n_new) if ch == A.CHOICE_PUSH
                 else pull_bytes(o["entries"][lv], o["words"][lv], n, n_new))
            p = prof.setdefault(KERNEL_OF_CHOICE[ch], {"ms": 0.0, "bytes": 0.0, "launches": 0})
            p["ms"] += float(o["kernel_ms"][lv])
            p["bytes"] += float(b)
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


def cpu_baseline(g, roots, gpu_levels):
    """The oracle's C + OpenMP restatement of ALBBA (reference C=256,
    sigma=16384, tasks = threads) on a bounded sample of the same roots."""
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rp, col = g.row_ptr, g.col
    t = time.time()
    s = O.build_sell(rp, col, 256, 16384)
    log(f"cpu baseline: oracle SELL build {time.time() - t:.1f}s")
    times, parity = [], True
    budget_s, spent = 20.0, 0.0
    for i, r in enumerate(roots):
        t0 = time.perf_counter()
        res = O.albba_bfs_on(s, g.n, int(r), tasks=threads)
        dt = time.perf_counter() - t0
        times.append(dt * 1e3)
        spent += dt
        if i in gpu_levels:
            parity &= bool(np.array_equal(res["level"], gpu_levels[i][0]) and
                           np.array_equal(res["parent"], gpu_levels[i][1]))
        log(f"cpu baseline: root {r} {dt * 1e3:.0f} ms")
        if spent > budget_s or i >= 7:
            break
    O.free_sell(s)
    import albba_amd as A

    gt = [A.gteps(g.nnz, ms) for ms in times]
    return {"value": round(A.harmonic_mean(gt), 4), "unit": "GTEPS", "cores": threads,
            "kind": "port",
            "sample": f"ALBBA level loop restated in C+OpenMP (oracle), {len(times)} of the same "
                      f"roots, same RMAT s24 graph, SELL C=256 sigma=16384, {threads} tasks; "
                      f"harmonic mean", "levels_match_gpu": parity}


def batch_specs(count, mix_seed):
    """Seeded instance mix: RMAT s14-20, ef 8-32, a 0.45-0.65 (b = c =
    (1 - a - 0.05) / 2); Erdos-Renyi n = 2^16..2^20, mean degree 4-16; square
    4-neighbour grids of side 256..2048 with 10% of the edges removed."""
    import albba_amd as A

    def u01(x):
        return (x >> 11) / float(1 << 53)

    specs = []
    for k in range(count):
        r = [A.rng_u64(mix_seed, 16 * k + i) for i in range(8)]
        kind = r[0] % 3
        if kind == 0:
            specs.append(("rmat", dict(scale=14 + r[1] % 7, edgefactor=8 + r[2] % 25,
                                       a=round(0.45 + 0.2 * u01(r[3]), 4),
                                       seed=r[4] % 1000003 + 1, perm_seed=r[5] % 1000003 + 1)))
        elif kind == 1:
            specs.append(("er", dict(n=1 << (16 + r[1] % 5), mean_degree=round(4 + 12 * u01(r[2]), 3),
                                     seed=r[4] % 1000003 + 1, perm_seed=r[5] % 1000003 + 1)))
        else:
            side = 256 + r[1] % 1793
            specs.append(("grid", dict(rows=side, cols=side, p_remove=0.1,
                                       seed=r[4] % 1000003 + 1, perm_seed=r[5] % 1000003 + 1)))
    return specs


def make_graph(kind, p):
    import albba_amd as A

    if kind == "rmat":
        b = (1.0 - p["a"] - 0.05) / 2.0
        return A.Graph.rmat(p["scale"], p["edgefactor"], p["a"], b, b, seed=p["seed"],
                            perm_seed=p["perm_seed"])
    if kind == "er":
        return A.Graph.er(p["n"], p["mean_degree"], seed=p["seed"], perm_seed=p["perm_seed"])
    return A.Graph.grid(p["rows"], p["cols"], p["p_remove"], seed=p["seed"], perm_seed=p["perm_seed"])


def run_batch_workload(args):
    """Batched multi-instance BFS: every instance in one cooperative launch.
    Sampled instances are checked bit for bit against the single-BFS path."""
    import albba_amd as A

    count, mix_seed = args.instances, 6
    specs = batch_specs(count, mix_seed)
    ctx = A.Context(0)
    sel = None
    if os.path.exists(SELECTOR):
        sel = A.Selector(ctx)
        sel.load(SELECTOR)
    # the selector is trained on RMAT levels; the batch defaults to the heuristic
    policy = A.SWITCH_LEARNED if sel is not None and args.policy == "learned" else A.SWITCH_HEURISTIC
    if policy == A.SWITCH_HEURISTIC:
        sel = None
    # layout per graph family from the autotuner, on the family's smallest instance
    t = time.time()
    tuned = {}
    for kind in ("rmat", "er", "grid"):
        idx = [i for i, (k, _) in enumerate(specs) if k == kind]
        if not idx:
            continue
        small = min(idx, key=lambda i: (specs[i][1].get("scale", 0), specs[i][1].get("n", 0),
                                        specs[i][1].get("rows", 0)))
        g = make_graph(*specs[small])
        roots = g.pick_roots(4, 5, largest_cc=(kind == "grid"))
        tuned[kind] = A.autotune(ctx, g, roots, policy=policy, selector=sel)["best"]
        log(f"autotune {kind}: C={tuned[kind]['C']} sigma={tuned[kind]['sigma']} "
            f"mapping={tuned[kind]['mapping']} ({tuned[kind]['ms']:.3f} ms/BFS)")
    log(f"autotuning {time.time() - t:.1f}s")
    t = time.time()
    dgs, roots, nnz, kinds, keep = [], [], [], [], {}
    for i, (kind, p) in enumerate(specs):
        g = make_graph(kind, p)
        cfg = tuned[kind]
        sigma = cfg["sigma"] if cfg["sigma"] in (cfg["C"], 256, 4096) else max(g.n, cfg["C"])
        sell = g.sellcs(cfg["C"], sigma, 1)
        dgs.append(A.DeviceGraph(ctx, sell))
        del sell
        r = g.pick_roots(1, 1000 + i, largest_cc=(kind == "grid"))
        roots.append(int(r[0]))
        nnz.append(g.nnz)
        kinds.append(kind)
        if i < 3 or i % 64 == 0:
            keep[i] = dgs[-1]  # checked against the single-BFS path below
        if (i + 1) % 32 == 0:
            log(f"  {i + 1}/{count} instances built ({time.time() - t:.1f}s)")
    mapping = [tuned[k]["mapping"] for k in kinds]
    b = A.Batch(ctx, dgs)
    out = b.run(roots, policy=policy, selector=sel, mapping=mapping)  # warm-up
    exact = True
    for i, dg in keep.items():
        lv, pa = b.results(i)
        o = dg.bfs(roots[i], policy=A.SWITCH_HEURISTIC)
        exact &= bool(np.array_equal(lv, o["level"]) and np.array_equal(pa, o["parent"]))
    del keep
    log(f"warm-up done, sampled instances bit-exact: {exact}")
    runs = []
    for s in range(args.steps):
        out = b.run(roots, policy=policy, selector=sel, mapping=mapping)
        runs.append(out)
        log(f"  step {s}: {out['total_ms']:.2f} ms")
    ms = float(np.mean([o["total_ms"] for o in runs]))
    total = float(np.sum(nnz))
    done = np.concatenate([o["done_ms"] for o in runs])
    serv = np.concatenate([o["ms"] for o in runs])
    line = {
        "metric": f"batched BFS total GTEPS, {count} instances (RMAT s14-20, ER 2^16-2^20, "
                  "grids 256^2-2048^2), one cooperative launch",
        "value": round(total / (ms * 1e6), 3), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32",
        "data": "synthetic: seeded mix (splitmix64 counter RNG), one root per instance",
        "config": {"workload": f"batch{count}_mix{mix_seed}", "instances": count,
                   "rmat": kinds.count("rmat"), "er": kinds.count("er"), "grid": kinds.count("grid"),
                   "nnz_stored_total": int(total), "policy": "learned" if sel else "heuristic",
                   "layouts": tuned},
        "latency_ms": {"p50": round(float(np.percentile(done, 50)), 3),
                       "p99": round(float(np.percentile(done, 99)), 3),
                       "service_p50": round(float(np.percentile(serv, 50)), 3),
                       "service_p99": round(float(np.percentile(serv, 99)), 3)},
        "levels_max": int(max(o["num_levels"].max() for o in runs)),
        "sampled_bit_exact": exact,
    }
    print(json.dumps(line), flush=True)
    b.close()
    for dg in dgs:
        dg.close()
    ctx.close()


def run_grid_workload(args):
    """High-diameter case: 4-neighbour grid of 4096 x 4096 vertices with 10%
    of the edges removed (seed 4, ids permuted with seed 2), 32 roots from
    the largest component (seed 5)."""
    import albba_amd as A

    t = time.time()
    g = A.Graph.grid(4096, 4096, 0.1, seed=4, perm_seed=2)
    roots = g.pick_roots(args.roots if args.roots != 64 else 32, 5, largest_cc=True)
    ctx = A.Context(0)
    dg = A.DeviceGraph(ctx, g.sellcs(64, 4096, 1))
    log(f"grid 4096^2: n={g.n} nnz={g.nnz} ({time.time() - t:.1f}s)")
    res = {}
    for name, cap in (("small_frontier_kernel", None), ("per_level_launches", 0)):
        dg.bfs(int(roots[0]), small_frontier_edges=cap, results=False)  # warm-up
        ms, levels = [], []
        for r in roots:
            o = dg.bfs(int(r), small_frontier_edges=cap, results=False, capacity=1 << 15)
            ms.append(o["bfs_ms"])
            levels.append(o["num_levels"])
        res[name] = {"hm_gteps": A.harmonic_mean([A.gteps(g.nnz, x) for x in ms]),
                     "mean_ms": float(np.mean(ms)), "levels_mean": float(np.mean(levels)),
                     "us_per_level": float(np.sum(ms) * 1e3 / np.sum(levels))}
        log(f"  {name}: {res[name]}")
        if args.steps <= 1:
            break
    R = res["small_frontier_kernel"]
    line = {
        "metric": "harmonic-mean GTEPS, 32 roots, 4096^2 grid (10% edges removed), heuristic switch",
        "value": round(R["hm_gteps"], 4), "unit": "GTEPS", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(R["mean_ms"] * len(roots), 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic: seeded grid, permuted ids",
        "config": {"workload": "grid4096_p0.1_32roots", "n": g.n, "nnz_stored": g.nnz,
                   "roots": len(roots), "layout": "SELL-C-sigma C=64 sigma=4096 pad=1"},
        "detail": {k: {kk: round(vv, 4) for kk, vv in v.items()} for k, v in res.items()},
    }
    print(json.dumps(line), flush=True)
    dg.close()
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
    ap.add_argument("--workload", choices=["rmat24", "grid", "batch", "load"], default="rmat24",
                    help="rmat24: the headline metric; grid: the 4096^2 high-diameter case; "
                         "batch: the batched multi-instance BFS; load: .cbd file to device graph")
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
    if len(policies) < 2:
        bit_exact = None  # nothing to compare against
    log(f"warmup done, policies bit-exact: {bit_exact}")

    results = {}
    for name, pol, s in policies:
        log(f"timing policy {name}")
        per_root, step_ms, comp = run_policy(dg, roots, pol, s, args.steps)
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

    main_name = "learned" if "learned" in results else "heuristic"
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
    roofline = {"bound": "hbm", "kernel": dom,
                "achieved": round(achieved / 1e9, 2), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic_from_profile(dom),
                "avg_launch_ms": round(P["ms"] / max(P["launches"], 1), 5),
                "avg_bytes_per_launch": round(P["bytes"] / max(P["launches"], 1)),
                "launches": P["launches"],
                "copy_kernel_gbps": round(2 * copy_bytes / (copy_ms * 1e-3) / 1e9, 1),
                "kernels": {k: {"launches": v["launches"],
                                "avg_ms": round(v["ms"] / max(v["launches"], 1), 5),
                                "gbps": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1)
                                if v["ms"] > 0 else None}
                            for k, v in sorted(prof.items())}}
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
        "learned_over_heuristic": (round(results["learned"]["wall_gteps"] /
                                         results["heuristic"]["wall_gteps"], 4)
                                   if len(results) == 2 else None),
        "hm_gteps_events": {k: round(v["hm_gteps_events"], 3) for k, v in results.items()},
        "hm_gteps_graph500": {k: round(v["hm_gteps_graph500"], 3) for k, v in results.items()},
        "bit_exact": bit_exact,
        "pcie_inclusive_gteps": round(pcie_gteps, 3),
        "sell_build": sell_build,
        "graph_generate_ms": {k: round(v, 1) for k, v in gen_report.items()},
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(g, roots, keep)
    print(json.dumps(line), flush=True)
    dg.close()
    ctx.close()


if __name__ == "__main__":
    main()
