"""The autotuner: every (C, sigma) candidate is measured, the fastest is
returned with the faster batched pull mapping, and the chosen layout gives
the oracle's levels and parents."""
import numpy as np
import pytest

import albba_amd as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_autotune_default_space(ctx):
    g = A.Graph.rmat(14, 16, seed=81, perm_seed=82)
    roots = g.pick_roots(4, 83)
    out = A.autotune(ctx, g, roots)
    res, best = out["results"], out["best"]
    assert {(r["C"], r["sigma"]) for r in res} == {(c, s) for c in (32, 64, 128)
                                                   for s in (c, 256, 4096, g.n)}
    assert all(r["status"] == 0 and r["ms"] > 0 for r in res)
    assert best["ms"] == min(r["ms"] for r in res)
    assert best["mapping"] in (A.MAP_LANE_ROW, A.MAP_WAVE_ROW) and best["batch_ms"] > 0
    dg = A.DeviceGraph(ctx, g.sellcs(best["C"], best["sigma"], 1))
    for r in roots:
        o = dg.bfs(int(r))
        lv, pa = O.bfs_queue(g.row_ptr, g.col, int(r))
        np.testing.assert_array_equal(o["level"], lv)
        np.testing.assert_array_equal(o["parent"], pa)
    dg.close()


def test_autotune_candidates_and_errors(ctx):
    g = A.Graph.grid(80, 90, 0.1, seed=84, perm_seed=85)
    roots = g.pick_roots(3, 86, largest_cc=True)
    out = A.autotune(ctx, g, roots, candidates=[(64, 4096), (48, 4096), (128, 256)])
    st = [r["status"] for r in out["results"]]
    assert st[0] == 0 and st[1] != 0 and st[2] == 0  # C = 48 is not a slice height
    assert (out["best"]["C"], out["best"]["sigma"]) in [(64, 4096), (128, 256)]
    with pytest.raises(A.AlbbaError):
        A.autotune(ctx, g, [g.n])
    with pytest.raises(A.AlbbaError):
        A.autotune(ctx, g, roots, candidates=[(48, 4096)])
