"""Seeded synthetic workloads shared by the benchmark and the tests.

RMAT24:        RMAT scale 24, edge factor 16, a/b/c = .57/.19/.19, seeds 1/2;
               64 roots (root seed 7).
GRID4096:      4-neighbour grid of 4096 x 4096 vertices, 10% of the edges
               removed (seed 4), ids permuted (seed 2); 32 roots from the
               largest connected component (root seed 5).
batch_specs(): the batched-BFS mix (RMAT s14-20, ER 2^16-2^20, grids
               256^2-2048^2), one root per instance.

The graphs come from the library's counter-based generators, on the host
(Graph.rmat / er / grid) or on the device (DeviceGraph.generate), which give
the same CSR bit for bit.
"""
from __future__ import annotations

import numpy as np

from . import (DeviceGraph, Graph, SWITCH_HEURISTIC, autotune, rng_u64)

RMAT24 = dict(scale=24, edgefactor=16, a=0.57, b=0.19, c=0.19, seed=1, perm_seed=2)
RMAT24_ROOTS = (64, 7)
GRID4096 = dict(rows=4096, cols=4096, p_remove=0.1, seed=4, perm_seed=2)
GRID4096_ROOTS = (32, 5)
BATCH_MIX_SEED = 6


def batch_specs(count: int, mix_seed: int = BATCH_MIX_SEED):
    """Seeded instance mix: RMAT s14-20, ef 8-32, a 0.45-0.65 (b = c =
    (1 - a - 0.05) / 2); Erdos-Renyi n = 2^16..2^20, mean degree 4-16; square
    4-neighbour grids of side 256..2048 with 10% of the edges removed.
    Returns [(kind, params)]."""

    def u01(x):
        return (x >> 11) / float(1 << 53)

    specs = []
    for k in range(count):
        r = [rng_u64(mix_seed, 16 * k + i) for i in range(8)]
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


def gen_args(kind: str, p: dict) -> dict:
    """Keyword arguments of DeviceGraph.generate for a spec."""
    if kind == "rmat":
        b = (1.0 - p["a"] - 0.05) / 2.0
        return dict(scale=p["scale"], edgefactor=p["edgefactor"], a=p["a"], b=b, c=b,
                    seed=p["seed"], perm_seed=p["perm_seed"])
    if kind == "er":
        return dict(n=p["n"], mean_degree=p["mean_degree"], seed=p["seed"], perm_seed=p["perm_seed"])
    return dict(rows=p["rows"], cols=p["cols"], p_remove=p["p_remove"], seed=p["seed"],
                perm_seed=p["perm_seed"])


def make_graph(kind: str, p: dict) -> Graph:
    """The host CSR of a spec."""
    a = gen_args(kind, p)
    if kind == "rmat":
        return Graph.rmat(a["scale"], a["edgefactor"], a["a"], a["b"], a["c"], seed=a["seed"],
                          perm_seed=a["perm_seed"])
    if kind == "er":
        return Graph.er(a["n"], a["mean_degree"], seed=a["seed"], perm_seed=a["perm_seed"])
    return Graph.grid(a["rows"], a["cols"], a["p_remove"], seed=a["seed"], perm_seed=a["perm_seed"])


def root_of(kind: str, g: Graph, index: int) -> int:
    """The root of batch instance `index` (grids: from the largest component)."""
    return int(g.pick_roots(1, 1000 + index, largest_cc=(kind == "grid"))[0])


def size_class(kind: str, p: dict) -> int:
    """Size class of a spec for the autotuner: log2 of its vertex count,
    rounded to the next even number (a class spans a factor of four)."""
    lg = int(np.ceil(np.log2(max(vertices(kind, p), 2))))
    return lg + (lg & 1)


def layout_sigma(cfg: dict, n: int) -> int:
    """The sigma of a tuned layout on a graph of n vertices (the autotuner's
    'whole graph' candidate is n)."""
    return cfg["sigma"] if cfg["sigma"] in (cfg["C"], 256, 4096) else max(n, cfg["C"])


def tune_layouts(ctx, specs, policy: int = SWITCH_HEURISTIC, selector=None, log=None):
    """One autotuner run per (family, size class) of the mix, on the class's
    smallest instance: {(kind, class): best config}."""
    groups = {}
    for i, (kind, p) in enumerate(specs):
        groups.setdefault((kind, size_class(kind, p)), []).append(i)
    tuned = {}
    for key in sorted(groups):
        kind = key[0]
        idx = groups[key]
        small = min(idx, key=lambda i: vertices(*specs[i]))
        g = make_graph(*specs[small])
        roots = g.pick_roots(4, 5, largest_cc=(kind == "grid"))
        tuned[key] = autotune(ctx, g, roots, policy=policy, selector=selector)["best"]
        if log:
            t = tuned[key]
            log(f"autotune {kind} class 2^{key[1]}: C={t['C']} sigma={t['sigma']} "
                f"mapping={t['mapping']} ({t['ms']:.3f} ms/BFS, {len(idx)} instances)")
    return tuned


def vertices(kind: str, p: dict) -> int:
    """Vertex count of a spec's graph."""
    return (1 << p["scale"]) if kind == "rmat" else (p["n"] if kind == "er" else p["rows"] * p["cols"])


def tuned_config(tuned: dict, kind: str, p: dict) -> dict:
    return tuned[(kind, size_class(kind, p))]


def generate_instance(ctx, kind: str, p: dict, cfg: dict, host_copy: bool = True):
    """Device graph of a spec in the tuned layout (generated on the device),
    with its host CSR when host_copy."""
    n = vertices(kind, p)
    return DeviceGraph.generate(ctx, kind, cfg["C"], layout_sigma(cfg, n), 1, host_copy=host_copy,
                                **gen_args(kind, p))
