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
        for i, r in enumerate(roots):
            o = dg.bfs(int(r), policy=policy, selector=sel, results=False,
                       script=None if scripts is None else scripts[i])
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


def run_batch_workload(args):
    """Batched multi-instance BFS: every instance in one launch.  Layouts from
    the autotuner per (family, size class); the heuristic and the learned
    direction policies on the same instances; sampled instances checked bit
    for bit against the single-BFS path (every instance is checked against
    the oracle by the GPU test of this workload)."""
    import albba_amd as A
    from albba_amd import workloads as W

    count = args.instances
    specs = W.batch_specs(count)
    ctx = A.Context(0)
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
    # a result that is not bit-exact carries no value and fails the run
    ok = bit_exact is not False and line.get("cpu_baseline", {}).get("levels_match_gpu", True)
    if not ok:
        line["value"] = None
        line["error"] = "levels or parents differ between policies or from the oracle"
    print(json.dumps(line), flush=True)
    if sel is not None:
        sel.close()
    dg.close()
    ctx.close()
    if not ok:
        sys.exit(1)


if __name__ == "__main__":
    main()
