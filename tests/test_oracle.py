"""Pins the CPU oracle before it is trusted as the checker.

 - SELL-C-sigma and run_bfs_ref frontier counts against fixtures written by
   the reference's own C code (tests/golden/make_golden.py), and against the
   live reference build when it is present;
 - BFS levels against scipy's unweighted shortest paths (textbook distances
   are unique) and a brute-force Python BFS on tiny graphs;
 - min-id parents against their closed form by brute force;
 - the validator against mistakes planted in oracle outputs.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as cg

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GG = np.load(os.path.join(HERE, "golden", "golden_graphs.npz"))
GRAPHS = sorted({k.split("__")[0] for k in GG.files})
SELL_CASES = [(256, 16384), (64, 4096), (64, 64), (32, 64), (128, 256), (16, 32)]
SELL_KEYS = ["row_order", "slice_widths", "vop_pointers", "vop_lengths", "slice_pointers",
             "column_indices", "max_index", "min_index"]


def graph(name):
    return GG[f"{name}__rp"], GG[f"{name}__col"]


def scipy_levels(rp, col, root):
    n = rp.size - 1
    W = sp.csr_matrix((np.ones(col.size), col, rp), shape=(n, n))
    # rows of W list in-neighbours, so out-edges are W^T
    d = cg.shortest_path(W.T.tocsr(), unweighted=True, directed=True, indices=root)
    return np.where(np.isinf(d), -1, d).astype(np.int32)


def brute_bfs(rp, col, root):
    n = rp.size - 1
    level = [-1] * n
    level[root] = 0
    changed = True
    while changed:
        changed = False
        for v in range(n):
            for k in range(rp[v], rp[v + 1]):
                u = col[k]
                if level[u] >= 0 and (level[v] < 0 or level[v] > level[u] + 1):
                    level[v] = level[u] + 1
                    changed = True
    parent = [-1] * n
    for v in range(n):
        if level[v] > 0:
            parent[v] = min(col[k] for k in range(rp[v], rp[v + 1]) if level[col[k]] == level[v] - 1)
    parent[root] = root
    return np.array(level, np.int32), np.array(parent, np.int32)


@pytest.mark.parametrize("name", GRAPHS)
@pytest.mark.parametrize("C_,sigma", SELL_CASES)
def test_oracle_sell_matches_reference_fixture(name, C_, sigma):
    rp, col = graph(name)
    got = O.sell_build(rp, col, C_, sigma)
    for k in SELL_KEYS:
        np.testing.assert_array_equal(got[k], GG[f"{name}__sell_{C_}_{sigma}__{k}"], err_msg=k)


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
def test_oracle_sell_matches_live_reference():
    import albba_amd as A

    for seed in range(3):
        g = A.Graph.rmat(11, 8 + 4 * seed, 0.5 + 0.05 * seed, 0.2, 0.2, seed=100 + seed,
                         perm_seed=7)
        rp, col = g.row_ptr.copy(), g.col.copy()
        for C_, sigma in [(256, 16384), (64, 128), (24, 96)]:
            ref = O.ref_sell(rp.astype(np.int32), col, C_, sigma)
            got = O.sell_build(rp, col, C_, sigma)
            for k in SELL_KEYS:
                np.testing.assert_array_equal(got[k], ref[k], err_msg=f"{k} C={C_}")


@pytest.mark.parametrize("name", GRAPHS)
def test_dense_ref_frontiers_match_reference_fixture(name):
    rp, col = graph(name)
    fr, lv = O.bfs_dense_ref(rp, col, 0)
    np.testing.assert_array_equal(fr, GG[f"{name}__bfs_ref_frontier"])
    assert 1 + int(fr[1:].sum()) == int(GG[f"{name}__bfs_ref_sum_vex"][0])
    # the frontier counts are the histogram of levels
    np.testing.assert_array_equal(np.bincount(lv[lv >= 0]), fr)


@pytest.mark.parametrize("name", GRAPHS)
def test_queue_bfs_levels_match_scipy_and_dense(name):
    rp, col = graph(name)
    n = rp.size - 1
    for root in sorted({0, n // 3, n - 1}):
        lv, pa = O.bfs_queue(rp, col, root)
        np.testing.assert_array_equal(lv, scipy_levels(rp, col, root))
        np.testing.assert_array_equal(O.bfs_dense_ref(rp, col, root)[1], lv)
        assert O.validate(rp, col, root, lv, pa) == 0


@pytest.mark.parametrize("name", ["rmat8", "grid16", "path50", "star200", "two_comp",
                                  "dir_general_real", "dir_mixed_case_comments"])
def test_queue_bfs_matches_brute_force(name):
    rp, col = graph(name)
    for root in (0, (rp.size - 1) // 2):
        lv, pa = O.bfs_queue(rp, col, root)
        blv, bpa = brute_bfs(rp, col, root)
        np.testing.assert_array_equal(lv, blv)
        np.testing.assert_array_equal(pa, bpa)


@pytest.mark.parametrize("name", GRAPHS)
@pytest.mark.parametrize("C_,sigma,tasks", [(256, 16384, 8), (64, 64, 3), (16, 32, 1),
                                            (32, 64, 5)])
def test_albba_restatement_levels_and_parents(name, C_, sigma, tasks):
    rp, col = graph(name)
    n = rp.size - 1
    for root in sorted({0, n // 2, n - 1}):
        lv, pa = O.bfs_queue(rp, col, root)
        r = O.albba_bfs(rp, col, root, C_, sigma, tasks)
        np.testing.assert_array_equal(r["level"], lv)
        np.testing.assert_array_equal(r["parent"], pa)
        np.testing.assert_array_equal(r["frontier"], np.bincount(lv[lv >= 0]))


def test_albba_restatement_covers_every_conversion():
    """The 0.04 n switch must exercise SpMV->SpMV, SpMV->SpMSpV, SpMSpV->SpMV
    and SpMSpV->SpMSpV (bfs/src/bfs_SELLCSv2.c:472-525) across these cases."""
    import albba_amd as A

    seen = set()
    for scale, ef in [(12, 4), (12, 16), (14, 8), (13, 32)]:
        g = A.Graph.rmat(scale, ef, seed=scale * ef, perm_seed=3)
        rp, col = g.row_ptr.copy(), g.col.copy()
        for root in g.pick_roots(4, 9):
            r = O.albba_bfs(rp, col, int(root), 256, 16384, 8)
            lv, pa = O.bfs_queue(rp, col, int(root))
            np.testing.assert_array_equal(r["level"], lv)
            np.testing.assert_array_equal(r["parent"], pa)
            m = r["mode"]
            seen |= {(int(a), int(b)) for a, b in zip(m[:-1], m[1:])}
    assert seen == {(0, 0), (0, 1), (1, 0), (1, 1)}


def test_f9_defect_loses_vertices_only_where_windows_cross_tasks():
    """The reference's dense2sparse (bfs/src/bfs_SELLCSv2.c:488-499,
    bfs/src/bfs_par_kernels.c:754-787) scans task t's id range but writes from
    offsets counted over task t's slice rows.  With sigma = C = 256 every
    window is one slice, the two agree and the level loop is exact; with
    sigma = 16384 windows cross task boundaries, frontier vertices are
    overwritten at a pull -> push switch, and the vertices only they reach
    come later or never.  The corrected restatement is exact on both."""
    import albba_amd as A

    g = A.Graph.rmat(14, 16, seed=224, perm_seed=3)
    rp, col = g.row_ptr.copy(), g.col.copy()
    n = rp.size - 1
    roots = [int(r) for r in g.pick_roots(16, 9)]
    for sigma in (256, 16384):
        s = O.build_sell(rp, col, 256, sigma)
        try:
            wrong = 0
            for root in roots:
                lv, pa = O.bfs_queue(rp, col, root)
                fixed = O.albba_bfs_on(s, n, root, 8)
                np.testing.assert_array_equal(fixed["level"], lv)
                np.testing.assert_array_equal(fixed["parent"], pa)
                r = O.albba_bfs_f9(s, n, root, 8)
                bad = r["level"] != lv
                # lost frontier vertices delay what they reach, never hasten it
                assert np.all((r["level"][bad] == -1) | (r["level"][bad] > lv[bad]))
                wrong += int(bad.sum())
            if sigma == 256:
                assert wrong == 0
            else:
                assert wrong > 0
        finally:
            O.free_sell(s)


def test_min_id_parent_closed_form():
    rng = np.random.default_rng(5)
    for trial in range(20):
        n = int(rng.integers(5, 40))
        edges = rng.integers(0, n, size=(int(rng.integers(n, 4 * n)), 2))
        rows = [set() for _ in range(n)]
        for u, v in edges:
            if u != v:
                rows[u].add(int(v))
                if trial % 2 == 0:
                    rows[v].add(int(u))  # half the cases directed
        rp = np.zeros(n + 1, np.int64)
        rp[1:] = np.cumsum([len(r) for r in rows])
        col = np.array([c for r in rows for c in sorted(r)], np.int32)
        root = int(rng.integers(0, n))
        lv, pa = O.bfs_queue(rp, col, root)
        blv, bpa = brute_bfs(rp, col, root)
        np.testing.assert_array_equal(lv, blv)
        np.testing.assert_array_equal(pa, bpa)
        assert O.validate(rp, col, root, lv, pa) == 0


def test_validator_rejects_planted_mistakes():
    rp, col = graph("rmat10")
    lv, pa = O.bfs_queue(rp, col, 0)
    assert O.validate(rp, col, 0, lv, pa) == 0
    reached = np.nonzero(lv > 1)[0]
    v = int(reached[0])
    # a wrong level
    bad = lv.copy()
    bad[v] += 1
    assert O.validate(rp, col, 0, bad, pa) != 0
    # a parent that is a valid neighbour one level up but not the smallest
    cands = [int(u) for u in col[rp[v]:rp[v + 1]] if lv[u] == lv[v] - 1]
    w = next(x for x in reached if len([u for u in col[rp[x]:rp[x + 1]] if lv[u] == lv[x] - 1]) > 1)
    cands = sorted(int(u) for u in col[rp[w]:rp[w + 1]] if lv[u] == lv[w] - 1)
    bad = pa.copy()
    bad[w] = cands[1]
    assert O.validate(rp, col, 0, lv, bad) == -6
    # a parent that is not a neighbour
    bad = pa.copy()
    non = [u for u in np.nonzero(lv == lv[v] - 1)[0] if u not in set(col[rp[v]:rp[v + 1]])]
    bad[v] = int(non[0])
    assert O.validate(rp, col, 0, lv, bad) == -3
    # an unreached vertex that has a reached in-neighbour
    bad = lv.copy()
    badp = pa.copy()
    leaf = int(np.argmax(lv))
    bad[leaf] = -1
    badp[leaf] = -1
    assert O.validate(rp, col, 0, bad, badp) != 0
    assert cands and len(cands) > 1


def test_mlp_oracle_gradient_matches_finite_differences():
    from oracle import mlp_oracle as M

    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.3, M.NPARAM)
    x = rng.random((17, M.NI))
    t = rng.normal(0, 1, (17, M.NO))
    _, g = M.loss_grad(p, x, t)
    idx = rng.choice(M.NPARAM, 120, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(M.NPARAM)
        e[i] = h
        fd = (M.loss(p + e, x, t) - M.loss(p - e, x, t)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
