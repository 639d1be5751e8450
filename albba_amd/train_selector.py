#!/usr/bin/env python3
"""Train the learned BFS direction selector on the GPU.

Training data come from the BFS itself: every level of a BFS run with the
all-choices policy (ALBBA_SWITCH_BEST) is executed once per kernel choice on
a snapshot of the state -- push SpMSpV, pull SpMV, pull over the unvisited
slice list, pull with frontier-word column bypass -- and the median kernel
time of each choice is recorded next to the 12 level features.  Graphs are
drawn from train_seed: RMAT of scale 18-23 with a in [0.45, 0.65]
(b = c = (1 - a - 0.05) / 2, d = 0.05) and edge factor 8-32, Erdos-Renyi
graphs of 2^18-2^22 vertices with mean degree 4-32, and square grids of
side 256-1024 with 10% of the edges removed; a BFS keeps at most 64 of its
levels (a sample of a grid's thousands).

Targets are relative log times, t'_c = log t_c - mean_k log t_k (each time
capped at 1000x the level's fastest), so the network ranks the choices; the
BFS runs the argmin.  The MLP 12 -> 32 -> 32 -> 4 is trained full-batch with
Adam by the library's own HIP forward/backward kernels.

Model selection: several networks (different initial weights) are trained
on the training graphs' levels; each runs the deployed level loop on
validation graphs that no benchmark or test uses (RMAT s22, s23 and s24 and
an Erdos-Renyi graph, with seeds of their own), and the one with the highest
geometric mean of its speedup over the heuristic switch across those graphs
is kept.  Each network's queue-listing ratio (see
albba_selector_set_queue_ratio), fixed before it is timed and stored with
its weights, is the threshold on frontier out-edges per vertex that best
separates, over its push levels on the validation graphs, those followed by
another push (which need the queue) from those followed by a pull (which do
not).  --validate-only applies the same selection to existing weight files
(an earlier selector among them).

    python -m albba_amd.train_selector --out albba_amd/selector.txt
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import albba_amd as A  # noqa: E402

# validation graphs: seeds used by no benchmark, test or training graph
VALIDATION = [("rmat", dict(scale=23, edgefactor=16, a=0.57, b=0.19, c=0.19, seed=101, perm_seed=102)),
              ("rmat", dict(scale=24, edgefactor=16, a=0.57, b=0.19, c=0.19, seed=201, perm_seed=202)),
              ("rmat", dict(scale=22, edgefactor=32, a=0.5, b=0.225, c=0.225, seed=301, perm_seed=302)),
              ("er", dict(n=1 << 22, mean_degree=16.0, seed=401, perm_seed=402))]


def u01(x):
    return (x >> 11) / float(1 << 53)


def graph_specs(train_seed: int, scales, per_scale: int, er: int, grids: int):
    """Deterministic list of (kind, generator arguments)."""
    specs = []
    k = 0
    for s in scales:
        for _ in range(per_scale):
            r = [A.rng_u64(train_seed, k * 8 + i) for i in range(4)]
            k += 1
            ef = 8 + int(r[0] % 25)                       # 8 .. 32
            a = round(0.45 + 0.20 * u01(r[1]), 4)         # 0.45 .. 0.65
            b = (1.0 - a - 0.05) / 2.0
            specs.append(("rmat", dict(scale=s, edgefactor=ef, a=a, b=b, c=b,
                                       seed=int(r[2] % 1000003) + 11, perm_seed=int(r[3] % 1000003) + 13)))
    for _ in range(er):
        r = [A.rng_u64(train_seed, k * 8 + i) for i in range(4)]
        k += 1
        specs.append(("er", dict(n=1 << (18 + int(r[0] % 5)), mean_degree=round(4 + 28 * u01(r[1]), 3),
                                 seed=int(r[2] % 1000003) + 11, perm_seed=int(r[3] % 1000003) + 13)))
    for _ in range(grids):
        r = [A.rng_u64(train_seed, k * 8 + i) for i in range(4)]
        k += 1
        side = 256 + int(r[0] % 769)
        specs.append(("grid", dict(rows=side, cols=side, p_remove=0.1,
                                   seed=int(r[2] % 1000003) + 11, perm_seed=int(r[3] % 1000003) + 13)))
    return specs


def heuristic_choice(frontier: int, n: int) -> int:
    return A.CHOICE_PULL_SB if frontier > int(n * 0.04) else A.CHOICE_PUSH


def collect(ctx, specs, roots_per_graph: int, reps: int, root_seed: int, max_levels: int, log=print):
    """Per-level features and per-choice kernel times.  A BFS with more than
    max_levels levels (a grid's thousands) contributes max_levels of them,
    drawn without replacement by a seeded generator, so that long BFS do not
    outweigh the graphs whose few levels carry the decisions."""
    X, T, H, G = [], [], [], []
    for gi, (kind, p) in enumerate(specs):
        t0 = time.time()
        dg, g = A.DeviceGraph.generate(ctx, kind, 64, 4096, 1, **p)
        for ri, r in enumerate(g.pick_roots(roots_per_graph, root_seed + gi, largest_cc=(kind == "grid"))):
            o = dg.bfs(int(r), policy=A.SWITCH_BEST, timing_reps=reps, results=False,
                       capacity=1 << 14)
            levels = range(o["num_levels"])
            if o["num_levels"] > max_levels:
                rng = np.random.default_rng([root_seed, gi, ri])
                levels = np.sort(rng.choice(o["num_levels"], max_levels, replace=False))
            for lv in levels:
                X.append(o["features"][lv])
                T.append(o["choice_ms"][lv])
                H.append(heuristic_choice(int(o["frontier"][lv]), g.n))
                G.append(gi)
        dg.close()
        log(f"graph {gi}: {kind} {p} n={g.n} nnz={g.nnz} samples={len(X)} ({time.time() - t0:.1f}s)")
    return (np.asarray(X, np.float32), np.asarray(T, np.float32), np.asarray(H, np.int32),
            np.asarray(G, np.int32))


def targets(choice_ms: np.ndarray) -> np.ndarray:
    t = np.maximum(choice_ms.astype(np.float64), 1e-4)
    t = np.minimum(t, 1000.0 * t.min(axis=1, keepdims=True))
    lt = np.log(t)
    return (lt - lt.mean(axis=1, keepdims=True)).astype(np.float32)


def decision_quality(pred_choice, choice_ms, heur_choice):
    """Total kernel time of the predicted choices vs the per-level best and
    the heuristic (ms, summed over levels)."""
    idx = np.arange(len(choice_ms))
    best = float(choice_ms.min(axis=1).sum())
    mine = float(choice_ms[idx, pred_choice].sum())
    heur = float(choice_ms[idx, heur_choice].sum())
    return {"best_ms": best, "model_ms": mine, "heuristic_ms": heur,
            "model_over_best": mine / best, "heuristic_over_best": heur / best,
            "accuracy": float(np.mean(pred_choice == choice_ms.argmin(axis=1)))}


def validate(ctx, graphs, selectors, roots: int, reps: int, log=print):
    """Each selector on every validation graph as deployed.  First its queue
    ratio: the threshold (queue_threshold) over its push levels on all the
    graphs (its choices do not depend on the ratio).  Then harmonic-mean
    GTEPS of the level loop over the roots (best of `reps` back-to-back
    passes, wall clock) for the heuristic, each selector and the per-level
    oracle choice (the all-choices policy's fastest choice of every level,
    replayed as a script)."""
    built = []
    for kind, p in graphs:
        dg, g = A.DeviceGraph.generate(ctx, kind, 64, 4096, 1, **p)
        built.append((kind, p, dg, g, g.pick_roots(roots, 17)))
    pairs_of = []
    for s in selectors:
        s.queue_ratio = 0.0
        pairs = []  # (frontier out-edges / n of a push level, next level pushed)
        for kind, p, dg, g, rts in built:
            for r in rts:
                o = dg.bfs(int(r), policy=A.SWITCH_LEARNED, selector=s, results=False)
                ch, ent = o["choice"], o["entries"]
                for lv in range(len(ch) - 1):
                    if ch[lv] == A.CHOICE_PUSH:
                        pairs.append((float(ent[lv]) / g.n, bool(ch[lv + 1] == A.CHOICE_PUSH)))
        s.queue_ratio = queue_threshold(pairs)
        pairs_of.append(pairs)
    out = {"heuristic": [], "oracle_choice": []}
    for k in range(len(selectors)):
        out[k] = []
    for kind, p, dg, g, rts in built:
        scripts = [dg.bfs(int(r), policy=A.SWITCH_BEST, timing_reps=3, results=False)["choice"].astype(np.int32)
                   for r in rts]
        pols = ([("heuristic", dict(policy=A.SWITCH_HEURISTIC))] +
                [(k, dict(policy=A.SWITCH_LEARNED, selector=s)) for k, s in enumerate(selectors)] +
                [("oracle_choice", dict(policy=A.SWITCH_SCRIPT, scripts=scripts))])
        for name, kw in pols:
            dg.bfs_many(rts[:2], **(kw if name != "oracle_choice" else dict(kw, scripts=scripts[:2])))
            ms = []
            for _ in range(reps):
                dg.ctx.synchronize()
                t0 = time.perf_counter()
                dg.bfs_many(rts, **kw)
                dg.ctx.synchronize()
                ms.append((time.perf_counter() - t0) * 1e3 / len(rts))
            out[name].append(A.gteps(g.nnz, float(np.min(ms))))
        log(f"validation {kind} {p}: " +
            ", ".join(f"{nm}={out[nm][-1]:.1f}" for nm in out))
    for kind, p, dg, g, rts in built:
        dg.close()
    hm = {nm: A.harmonic_mean(v) for nm, v in out.items()}
    return hm, out, pairs_of


def queue_threshold(pairs):
    """The ratio t minimizing the push levels put on the wrong side: followed
    by a push with out-edges/n above t (queue missing), or by a pull at or
    below t (queue listed for nothing).  Ties go to the larger t."""
    if not pairs:
        return 0.0
    cand = sorted({0.0} | {e for e, _ in pairs})
    best_t, best_err = 0.0, None
    for t in cand:
        err = sum(1 for e, nxt in pairs if (nxt and e > t) or (not nxt and e <= t))
        if best_err is None or err <= best_err:
            best_t, best_err = t, err
    return best_t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(__file__), "selector.txt"))
    ap.add_argument("--data-out", default=None, help="save the dataset (.npz)")
    ap.add_argument("--data-in", default=None, help="train from a saved dataset")
    ap.add_argument("--scales", default="18,19,20,21,22,23")
    ap.add_argument("--per-scale", type=int, default=5)
    ap.add_argument("--er", type=int, default=10)
    ap.add_argument("--grids", type=int, default=3)
    ap.add_argument("--roots", type=int, default=10)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--train-seed", type=int, default=8)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--candidates", type=int, default=12, help="networks trained (init seeds 1..k)")
    ap.add_argument("--max-levels", type=int, default=64, help="levels kept per BFS (sampled past it)")
    ap.add_argument("--incumbent", nargs="*", default=[], metavar="FILE",
                    help="existing selector files validated beside the new networks")
    ap.add_argument("--val-roots", type=int, default=16)
    ap.add_argument("--candidates-dir", default=None, help="also save every candidate's weights here")
    ap.add_argument("--validate-only", nargs="+", default=None, metavar="FILE",
                    help="train nothing: choose among these selector files by the same validation")
    args = ap.parse_args()

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    ctx = A.Context(0)
    if args.validate_only:
        sels = []
        for path in args.validate_only:
            sl = A.Selector(ctx)
            sl.load(path)
            sels.append(sl)
        report = {"candidates": [{"file": path} for path in args.validate_only]}
        choose(args, sels, report, log, lambda k: args.validate_only[k])
        return
    if args.data_in:
        d = np.load(args.data_in)
        X, Tms, H, G = d["X"], d["T"], d["H"], d["G"]
    else:
        specs = graph_specs(args.train_seed, [int(s) for s in args.scales.split(",")],
                            args.per_scale, args.er, args.grids)
        X, Tms, H, G = collect(ctx, specs, args.roots, args.reps, 1000 + args.train_seed, args.max_levels, log)
        if args.data_out:
            np.savez_compressed(args.data_out, X=X, T=Tms, H=H, G=G)
    Y = targets(Tms)
    # held-out levels for the decision-quality report: every fifth graph
    val = (G % 5) == 4
    tr = ~val
    report = {"train_samples": int(tr.sum()), "heldout_samples": int(val.sum()), "candidates": []}
    sels = []
    for k in range(1, args.candidates + 1):
        sel = A.Selector(ctx, seed=k)
        t0 = time.time()
        loss = sel.train(X[tr], Y[tr], steps=args.steps, lr=args.lr)
        q = decision_quality(sel.predict(X[val]).argmin(axis=1), Tms[val], H[val])
        log(f"candidate {k}: loss {loss:.5f} ({time.time() - t0:.2f}s), held-out levels "
            f"{q['model_over_best']:.3f}x best, heuristic {q['heuristic_over_best']:.3f}x")
        report["candidates"].append({"init_seed": k, "train_loss": loss, "heldout": q})
        sels.append(sel)
    ntrained = len(sels)
    for path in args.incumbent:
        sl = A.Selector(ctx)
        sl.load(path)
        q = decision_quality(sl.predict(X[val]).argmin(axis=1), Tms[val], H[val])
        report["candidates"].append({"file": path, "heldout": q})
        sels.append(sl)
    choose(args, sels, report, log,
           lambda k: f"candidate {k + 1}" if k < ntrained else args.incumbent[k - ntrained])


def choose(args, sels, report, log, name):
    """Validate the selectors, keep the one with the highest geometric-mean
    speedup over the heuristic across the validation graphs (with its queue
    ratio) in args.out, and print the report."""
    hm, per_graph, pairs_of = validate(sels[0].ctx, VALIDATION, sels, args.val_roots, 3, log)
    heur = np.asarray(per_graph["heuristic"])
    score = {k: float(np.exp(np.mean(np.log(np.asarray(per_graph[k]) / heur)))) for k in range(len(sels))}
    best = max(range(len(sels)), key=lambda k: score[k])
    orc = np.asarray(per_graph["oracle_choice"])
    for k, c in enumerate(report["candidates"]):
        c["validation_speedup_geomean"] = score[k]
        c["validation_hm_gteps"] = hm[k]
        c["validation_gteps"] = per_graph[k]
        c["validation_over_oracle_choice"] = (np.asarray(per_graph[k]) / orc).tolist()
        c["queue_ratio"] = sels[k].queue_ratio
    report["validation_heuristic_gteps"] = per_graph["heuristic"]
    report["validation_oracle_choice_gteps"] = per_graph["oracle_choice"]
    report["validation_graphs"] = [f"{kind} {p}" for kind, p in VALIDATION]
    report["chosen"] = name(best)
    pairs = pairs_of[best]
    report["queue_ratio"] = sels[best].queue_ratio
    report["queue_ratio_levels"] = {"push_then_push": sum(1 for _, nxt in pairs if nxt),
                                    "push_then_pull": sum(1 for _, nxt in pairs if not nxt)}
    if args.candidates_dir:
        os.makedirs(args.candidates_dir, exist_ok=True)
        for k, sl in enumerate(sels):
            sl.save(os.path.join(args.candidates_dir, f"selector_{k + 1}.txt"))
    sels[best].save(args.out)
    print(json.dumps(report, indent=1))
    log(f"chose {name(best)} (validation speedup over the heuristic {score[best]:.3f}, "
        f"geometric mean); queue ratio {sels[best].queue_ratio:.3g}; wrote {args.out}")


if __name__ == "__main__":
    main()
