"""Batched BFS (one cooperative launch over many instances) against the
oracle: every instance's levels bit for bit, its min-id parents bit for bit
and as a valid BFS tree, for both pull mappings, the heuristic and learned
direction policies, more instances than workgroups (teams run several
instances in turn) and a graph shared by several instances."""
import os

import numpy as np
import pytest

import albba_amd as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def mixed_graphs():
    return [A.Graph.rmat(11, 8, seed=61, perm_seed=62),
            A.Graph.rmat(13, 16, 0.6, 0.175, 0.175, seed=63, perm_seed=64),
            A.Graph.rmat(12, 32, 0.45, 0.25, 0.25, seed=65, perm_seed=66),
            A.Graph.er(5000, 4.0, seed=67, perm_seed=68),
            A.Graph.er(20000, 12.0, seed=69, perm_seed=70),
            A.Graph.grid(60, 70, 0.1, seed=71, perm_seed=72),
            A.Graph.grid(300, 20, 0.05, seed=73, perm_seed=74)]


@pytest.fixture(scope="module")
def batch_setup(ctx):
    graphs = mixed_graphs()
    dgs = [A.DeviceGraph(ctx, g.sellcs(64, 4096, 1)) for g in graphs]
    # two more instances on graph 0 and one on graph 5, with other roots
    inst = list(range(len(graphs))) + [0, 0, 5]
    roots = []
    for k, gi in enumerate(inst):
        r = graphs[gi].pick_roots(1, 100 + k, largest_cc=True)
        roots.append(int(r[0]))
    roots[-1] = 0  # a root outside the largest component is allowed too
    b = A.Batch(ctx, [dgs[gi] for gi in inst])
    yield graphs, inst, np.array(roots, np.int32), b
    b.close()
    for dg in dgs:
        dg.close()


def check_batch(graphs, inst, roots, b, out):
    for i, gi in enumerate(inst):
        g = graphs[gi]
        lv, pa = b.results(i)
        rlv, rpa = O.bfs_queue(g.row_ptr, g.col, int(roots[i]))
        np.testing.assert_array_equal(lv, rlv, err_msg=f"instance {i}")
        np.testing.assert_array_equal(pa, rpa, err_msg=f"instance {i}")
        assert O.validate(g.row_ptr, g.col, int(roots[i]), lv, pa) == 0
        assert out["error"][i] == 0
        assert out["reached"][i] == int((rlv >= 0).sum())
        assert out["num_levels"][i] == int(rlv.max()) + 1
        deg = np.diff(g.row_ptr)
        assert out["edges_in_component"][i] == int(deg[rlv >= 0].sum())
        assert out["push_levels"][i] + out["pull_levels"][i] == out["num_levels"][i]
        assert out["ms"][i] > 0 and out["done_ms"][i] >= out["ms"][i] * 0.999


@pytest.mark.parametrize("mapping", [A.MAP_LANE_ROW, A.MAP_WAVE_ROW])
@pytest.mark.parametrize("workgroups", [0, 3, 16])
def test_batch_matches_oracle(batch_setup, mapping, workgroups):
    graphs, inst, roots, b = batch_setup
    out = b.run(roots, mapping=mapping, workgroups=workgroups)
    check_batch(graphs, inst, roots, b, out)
    if workgroups == 3:
        # ten instances on three one-workgroup teams
        assert (out["team_size"] == 1).all() and len(set(out["team"].tolist())) == 3
    # a second run on the same scratch (keys and bitmaps were restored)
    out2 = b.run(roots[::-1].copy()[::-1], mapping=mapping, workgroups=workgroups)
    check_batch(graphs, inst, roots, b, out2)


def test_batch_repeated_runs(batch_setup):
    """Eight runs in a row, every instance checked each time.  A regression
    test for a race between a team's workgroups: one past a level's barrier
    appending for the next level while another still read the level's
    count; it showed as wrong levels on one machine in every run."""
    graphs, inst, roots, b = batch_setup
    for rep in range(8):
        check_batch(graphs, inst, roots, b, b.run(roots, mapping=rep % 2))


def test_batch_learned_policy_and_mixed_mappings(ctx, batch_setup):
    graphs, inst, roots, b = batch_setup
    sel = A.Selector(ctx)
    sel.load(os.path.join(os.path.dirname(A.__file__), "selector.txt"))
    mapping = np.arange(len(inst)) % 2
    out = b.run(roots, policy=A.SWITCH_LEARNED, selector=sel, mapping=mapping)
    check_batch(graphs, inst, roots, b, out)
    assert out["pull_levels"].sum() > 0 and out["push_levels"].sum() > 0
    sel.close()


def test_batch_errors(ctx, batch_setup):
    graphs, inst, roots, b = batch_setup
    bad = roots.copy()
    bad[2] = graphs[inst[2]].n
    with pytest.raises(A.AlbbaError):
        b.run(bad)
    with pytest.raises(A.AlbbaError):
        b.run(roots, policy=A.SWITCH_FIXED)
    with pytest.raises(A.AlbbaError):
        b.run(roots, policy=A.SWITCH_LEARNED)  # no selector
    with pytest.raises(A.AlbbaError):
        b.run(roots, mapping=7)
    with pytest.raises(ValueError):
        b.run(roots[:3])
    # the batch still works after rejected calls
    check_batch(graphs, inst, roots, b, b.run(roots))


def test_batch_team_sizes_from_previous_run(ctx):
    """From the second run on, teams are sized from the previous run's level
    counts and times: a long narrow grid (about 2 000 levels, few entries)
    next to RMAT and ER instances gets a bigger team than its share of the
    entries gave it, the teams never use more than the workgroups, every
    run stays exact, and forget() returns to the share split."""
    graphs = [A.Graph.grid(2000, 4, 0.0, seed=81, perm_seed=82)] + \
             [A.Graph.rmat(14, 16, seed=83 + 2 * k, perm_seed=84 + 2 * k) for k in range(6)] + \
             [A.Graph.er(30000, 8.0, seed=97, perm_seed=98)]
    dgs = [A.DeviceGraph(ctx, g.sellcs(64, 4096, 1)) for g in graphs]
    roots = np.array([int(g.pick_roots(1, 200 + k, largest_cc=True)[0]) for k, g in enumerate(graphs)],
                     np.int32)
    inst = list(range(len(graphs)))
    b = A.Batch(ctx, dgs)
    W = 64
    try:
        runs = []
        for k in range(3):
            if k == 2:
                b.forget()
            out = b.run(roots, workgroups=W)
            check_batch(graphs, inst, roots, b, out)
            teams = {int(t): int(s) for t, s in zip(out["team"], out["team_size"])}
            assert sum(teams.values()) <= W
            runs.append(out)
        assert runs[0]["num_levels"][0] > 1500
        assert runs[1]["team_size"][0] > runs[0]["team_size"][0]
        np.testing.assert_array_equal(runs[2]["team_size"], runs[0]["team_size"])
    finally:
        b.close()
        for dg in dgs:
            dg.close()
