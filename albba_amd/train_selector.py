#!/usr/bin/env python3
"""Train the learned BFS direction selector on the GPU.

Training data come from the BFS itself: every level of a BFS run with the
all-choices policy (ALBBA_SWITCH_BEST) is executed once per kernel choice on
a snapshot of the state -- push SpMSpV, pull SpMV, pull over the unvisited
slice list, pull with frontier-word column bypass -- and the median kernel
time of each choice is recorded next to the 12 level features.  Graphs are
RMAT of scale 18-22 with a in [0.45, 0.65] (b = c = (1 - a - 0.05) / 2,
d = 0.05) and edge factor 8-32, drawn from train_seed.

Targets are relative log times, t'_c = log t_c - mean_k log t_k (each time
capped at 1000x the level's fastest), so the network ranks the choices; the
BFS runs the argmin.  The MLP 12 -> 32 -> 32 -> 4 is trained full-batch with
Adam by the library's own HIP forward/backward kernels.

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


def graph_specs(train_seed: int, scales, per_scale: int):
    """Deterministic (scale, edgefactor, a, graph_seed, perm_seed) list."""
    specs = []
    k = 0
    for s in scales:
        for _ in range(per_scale):
            r = [A.rng_u64(train_seed, k * 8 + i) for i in range(4)]
            k += 1
            ef = 8 + int(r[0] % 25)                       # 8 .. 32
            a = 0.45 + 0.20 * (r[1] >> 11) / float(1 << 53)  # 0.45 .. 0.65
            specs.append((s, ef, round(a, 4), int(r[2] % 1000003) + 11, int(r[3] % 1000003) + 13))
    return specs


def heuristic_choice(frontier: int, n: int) -> int:
    return A.CHOICE_PULL_SB if frontier > int(n * 0.04) else A.CHOICE_PUSH


def collect(ctx, specs, roots_per_graph: int, reps: int, root_seed: int, log=print):
    X, T, H, G = [], [], [], []
    for gi, (scale, ef, a, gseed, pseed) in enumerate(specs):
        b = c = (1.0 - a - 0.05) / 2.0
        t0 = time.time()
        g = A.Graph.rmat(scale, ef, a, b, c, seed=gseed, perm_seed=pseed)
        dg = A.DeviceGraph(ctx, g.sellcs(64, 4096, 1))
        for r in g.pick_roots(roots_per_graph, root_seed + gi):
            o = dg.bfs(int(r), policy=A.SWITCH_BEST, timing_reps=reps, results=False)
            for lv in range(o["num_levels"]):
                X.append(o["features"][lv])
                T.append(o["choice_ms"][lv])
                H.append(heuristic_choice(int(o["frontier"][lv]), g.n))
                G.append(gi)
        dg.close()
        log(f"graph {gi}: s{scale} ef{ef} a={a} n={g.n} nnz={g.nnz} "
            f"samples={len(X)} ({time.time() - t0:.1f}s)")
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(__file__), "selector.txt"))
    ap.add_argument("--data-out", default=None, help="save the dataset (.npz)")
    ap.add_argument("--data-in", default=None, help="train from a saved dataset")
    ap.add_argument("--scales", default="18,19,20,21,22")
    ap.add_argument("--per-scale", type=int, default=5)
    ap.add_argument("--roots", type=int, default=10)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--train-seed", type=int, default=8)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--init-seed", type=int, default=1)
    args = ap.parse_args()

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    ctx = A.Context(0)
    if args.data_in:
        d = np.load(args.data_in)
        X, Tms, H, G = d["X"], d["T"], d["H"], d["G"]
    else:
        specs = graph_specs(args.train_seed, [int(s) for s in args.scales.split(",")],
                            args.per_scale)
        X, Tms, H, G = collect(ctx, specs, args.roots, args.reps, 1000 + args.train_seed, log)
        if args.data_out:
            np.savez_compressed(args.data_out, X=X, T=Tms, H=H, G=G)
    Y = targets(Tms)
    # hold out every fifth graph for validation
    val = (G % 5) == 4
    tr = ~val
    sel = A.Selector(ctx, seed=args.init_seed)
    t0 = time.time()
    loss = sel.train(X[tr], Y[tr], steps=args.steps, lr=args.lr)
    log(f"trained {args.steps} Adam steps on {tr.sum()} samples in {time.time() - t0:.2f}s, "
        f"loss {loss:.5f}")
    report = {"train_samples": int(tr.sum()), "val_samples": int(val.sum()), "train_loss": loss}
    for name, m in (("train", tr), ("val", val)):
        if m.sum() == 0:
            continue
        pred = sel.predict(X[m]).argmin(axis=1)
        report[name] = decision_quality(pred, Tms[m], H[m])
        report[name + "_loss"] = sel.loss_grad(X[m], Y[m])[0]
    # final model: train on every sample
    sel_all = A.Selector(ctx, seed=args.init_seed)
    report["all_loss"] = sel_all.train(X, Y, steps=args.steps, lr=args.lr)
    report["all"] = decision_quality(sel_all.predict(X).argmin(axis=1), Tms, H)
    sel_all.save(args.out)
    print(json.dumps(report, indent=1))
    log(f"wrote {args.out}")


if __name__ == "__main__":
    main()
