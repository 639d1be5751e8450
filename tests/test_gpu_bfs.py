"""GPU parity: the HIP BFS through the C ABI against the oracle.

Levels are compared bit for bit; parents bit for bit under the min-id rule
and checked as a BFS tree; per-level frontier sizes must equal the level
histogram.  Every kernel choice (push, pull, pull over the unvisited-slice
list, pull with column bypass), every switch policy, C in {16,32,64,128} and
pad in {1,16,32} must give the same result.
"""
import os

import numpy as np
import pytest

import albba_amd as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GG = np.load(os.path.join(HERE, "golden", "golden_graphs.npz"))
GOLDEN = sorted({k.split("__")[0] for k in GG.files})
POLICIES = [("heuristic", A.SWITCH_HEURISTIC, 0), ("push", A.SWITCH_FIXED, A.CHOICE_PUSH),
            ("pull", A.SWITCH_FIXED, A.CHOICE_PULL), ("pull_sb", A.SWITCH_FIXED, A.CHOICE_PULL_SB),
            ("pull_cb", A.SWITCH_FIXED, A.CHOICE_PULL_CB), ("best", A.SWITCH_BEST, 0)]


def modes_as_choices(mode):
    """The oracle's per-level mode (1: SpMV pull, 0: SpMSpV) as the kernel
    choice the heuristic policy takes for it (the slice-bypass pull, or the
    push)."""
    return np.where(np.asarray(mode) == 1, A.CHOICE_PULL_SB, A.CHOICE_PUSH)


def check(dg, rp, col, root, policy=A.SWITCH_HEURISTIC, fixed=0, **kw):
    o = dg.bfs(int(root), policy=policy, fixed_choice=fixed, **kw)
    lv, pa = O.bfs_queue(rp, col, int(root))
    np.testing.assert_array_equal(o["level"], lv)
    np.testing.assert_array_equal(o["parent"], pa)
    np.testing.assert_array_equal(o["frontier"], np.bincount(lv[lv >= 0]))
    assert o["reached"] == int((lv >= 0).sum())
    # out-edges of the reached vertices (a vertex's push out-degree is the
    # number of rows listing it; equal to its row length when symmetric)
    deg = np.bincount(np.asarray(col), minlength=len(lv))
    assert o["edges_in_component"] == int(deg[lv >= 0].sum())
    assert O.validate(rp, col, int(root), o["level"], o["parent"]) == 0
    return o


@pytest.mark.parametrize("name", GOLDEN)
@pytest.mark.parametrize("C_,pad", [(64, 1), (64, 32), (16, 16), (32, 1), (128, 16)])
def test_small_graphs_all_policies(ctx, name, C_, pad):
    rp, col = GG[f"{name}__rp"], GG[f"{name}__col"]
    g = A.Graph.from_arrays(rp, col, symmetric=not name.startswith("dir_"))
    n = g.n
    dg = A.DeviceGraph(ctx, g.sellcs(C_, C_ if n < 4096 else 4096, pad))
    for root in sorted({0, n // 2, n - 1}):
        for pname, pol, fc in POLICIES:
            check(dg, rp, col, root, pol, fc)
    dg.close()


@pytest.mark.parametrize("kind", ["rmat", "er", "grid"])
@pytest.mark.parametrize("C_", [16, 32, 64, 128])
def test_medium_graphs_every_choice(ctx, kind, C_):
    g = {"rmat": lambda: A.Graph.rmat(15, 16, seed=31, perm_seed=32),
         "er": lambda: A.Graph.er(40000, 6.0, seed=33, perm_seed=34),
         "grid": lambda: A.Graph.grid(150, 170, 0.1, seed=35, perm_seed=36)}[kind]()
    rp, col = g.row_ptr.copy(), g.col.copy()
    dg = A.DeviceGraph(ctx, g.sellcs(C_, 4096, 1))
    for root in g.pick_roots(3, 5):
        for pname, pol, fc in POLICIES[:5]:
            check(dg, rp, col, root, pol, fc)
    dg.close()


def test_random_scripts_mix_every_transition(ctx):
    """Random per-level choices exercise every conversion between the
    frontier queue, the bitmap and the unvisited-slice list."""
    g = A.Graph.rmat(14, 16, seed=41, perm_seed=42)
    rp, col = g.row_ptr.copy(), g.col.copy()
    dg = A.DeviceGraph(ctx, g.sellcs(64, 4096, 1))
    rng = np.random.default_rng(7)
    for root in g.pick_roots(3, 11):
        for _ in range(12):
            check(dg, rp, col, root, A.SWITCH_SCRIPT, script=rng.integers(0, 4, 16))
    dg.close()


def test_push_chunks_of_degree_one_vertices(ctx):
    """Frontiers made of long runs of degree-1 vertices: a 64-edge chunk of
    the push expansion spans 64 vertices (regression for the chunk walk)."""
    n = 6000
    edges = [(0, i) for i in range(1, 4001)] + [(i, 4000 + i) for i in range(1, 2000)]
    rows = [[] for _ in range(n)]
    for u, v in edges:
        rows[u].append(v)
        rows[v].append(u)
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    col = np.array([c for r in rows for c in sorted(r)], np.int32)
    g = A.Graph.from_arrays(rp, col, symmetric=True)
    dg = A.DeviceGraph(ctx, g.sellcs(64, 64, 1))
    for root in (0, 1, 4001, 2500):
        check(dg, rp, col, root, A.SWITCH_FIXED, A.CHOICE_PUSH)
        check(dg, rp, col, root, A.SWITCH_HEURISTIC)
    dg.close()


def test_isolated_root_and_errors(ctx):
    rp = np.array([0, 1, 2, 2, 2], np.int64)
    col = np.array([1, 0], np.int32)
    g = A.Graph.from_arrays(rp, col, symmetric=True)
    dg = A.DeviceGraph(ctx, g.sellcs(16, 16, 1))
    o = check(dg, rp, col, 3, A.SWITCH_FIXED, A.CHOICE_PUSH)
    assert o["reached"] == 1 and o["num_levels"] == 1
    check(dg, rp, col, 2, A.SWITCH_FIXED, A.CHOICE_PULL)
    for bad in (-1, 4, 10 ** 6):
        with pytest.raises(A.AlbbaError) as e:
            dg.bfs(bad)
        assert e.value.code == -1
    with pytest.raises(A.AlbbaError):
        dg.bfs(0, policy=9)
    with pytest.raises(A.AlbbaError):
        dg.bfs(0, policy=A.SWITCH_FIXED, fixed_choice=7)
    with pytest.raises(A.AlbbaError):
        dg.bfs(0, policy=A.SWITCH_LEARNED)  # no selector given
    dg.close()


@pytest.mark.parametrize("name", GOLDEN)
def test_run_custom_bfs_test_matches_reference_frontiers(name):
    """bfs/src/bfs.h:13 entry point: frontier sizes per iteration from vertex
    0 equal the reference's run_bfs_ref counts (fixture)."""
    rp, col = GG[f"{name}__rp"], GG[f"{name}__col"]
    out = A.run_custom_bfs_test(rp.size - 1, rp.astype(np.int32), col)
    np.testing.assert_array_equal(out["frontier"], GG[f"{name}__bfs_ref_frontier"])
    assert out["num_vex"] == int(GG[f"{name}__bfs_ref_sum_vex"][0])
    assert out["nnzx"] == 0 and out["bfs_time"] > 0


def test_config0_rmat16_sixteen_roots(ctx):
    """RMAT s16, ef 16, 16 seeded non-isolated roots: every policy and both
    the device layout and the reference layout's C agree with the oracle."""
    g = A.Graph.rmat(16, 16, 0.57, 0.19, 0.19, seed=1, perm_seed=2)
    rp, col = g.row_ptr.copy(), g.col.copy()
    roots = g.pick_roots(16, 3)
    assert len(roots) == 16
    for C_, sigma, pad in [(64, 4096, 1), (128, 16384, 32)]:
        dg = A.DeviceGraph(ctx, g.sellcs(C_, sigma, pad))
        for root in roots:
            o = check(dg, rp, col, root)
            ref = O.albba_bfs(rp, col, int(root), 256, 16384, 8)
            np.testing.assert_array_equal(o["frontier"], ref["frontier"])
            # the heuristic's direction per level is the reference's mode
            np.testing.assert_array_equal(o["choice"], modes_as_choices(ref["mode"]))
        dg.close()


def test_config1_rmat22_levels_bit_exact(ctx):
    """RMAT s22, ef 16, C=64 sigma=4096: 64 seeded roots, levels and parents
    bit-exact against the oracle's ALBBA restatement."""
    g = A.Graph.rmat(22, 16, 0.57, 0.19, 0.19, seed=1, perm_seed=2)
    rp, col = g.row_ptr, g.col
    roots = g.pick_roots(64, 3)
    dg = A.DeviceGraph(ctx, g.sellcs(64, 4096, 1))
    s = O.build_sell(rp, col, 256, 16384)
    try:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        for i, root in enumerate(roots):
            o = dg.bfs(int(root))
            ref = O.albba_bfs_on(s, g.n, int(root), tasks=threads)
            np.testing.assert_array_equal(o["level"], ref["level"])
            np.testing.assert_array_equal(o["parent"], ref["parent"])
            np.testing.assert_array_equal(o["choice"], modes_as_choices(ref["mode"]))
            if i < 2:
                assert O.validate(rp, col, int(root), o["level"], o["parent"]) == 0
    finally:
        O.free_sell(s)
        dg.close()


def test_heuristic_threshold_boundary(ctx):
    """Frontiers of exactly (int)(0.04 n) and one more vertex: the first
    stays a push, the second pulls (bfs/src/bfs_SELLCSv2.c:282-290, 302:
    pull iff nnzx > threshold).  n = 1000, threshold 40; level sizes 1, 40,
    41, 40, 41, 40, 41 and isolated vertices after them."""
    n, sizes = 1000, [1, 40, 41, 40, 41, 40, 41]
    rows = [[] for _ in range(n)]
    levels, nxt = [], 0
    for s in sizes:
        levels.append(list(range(nxt, nxt + s)))
        nxt += s
    for a, b in zip(levels, levels[1:]):
        for k, v in enumerate(b):
            for u in {a[k % len(a)], a[(3 * k + 1) % len(a)]}:
                rows[u].append(v)
                rows[v].append(u)
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    col = np.array([c for r in rows for c in sorted(r)], np.int32)
    g = A.Graph.from_arrays(rp, col, symmetric=True)
    ref = O.albba_bfs(rp, col, 0, 256, 16384, 4)
    np.testing.assert_array_equal(ref["frontier"], sizes)
    np.testing.assert_array_equal(ref["mode"], [0, 0, 1, 0, 1, 0, 1])
    for C_ in (16, 64):
        dg = A.DeviceGraph(ctx, g.sellcs(C_, 256, 1))
        o = check(dg, rp, col, 0)
        np.testing.assert_array_equal(o["choice"], modes_as_choices(ref["mode"]))
        dg.close()


def test_grid_many_levels(ctx):
    """4-neighbour grid with 10% of edges removed (the high-diameter case at
    a reduced size): hundreds of levels with small frontiers."""
    g = A.Graph.grid(400, 400, 0.1, seed=4, perm_seed=2)
    rp, col = g.row_ptr.copy(), g.col.copy()
    dg = A.DeviceGraph(ctx, g.sellcs(64, 4096, 1))
    for root in g.pick_roots(3, 5, largest_cc=True):
        o = check(dg, rp, col, root, capacity=1 << 14)
        assert o["num_levels"] > 300
        check(dg, rp, col, root, A.SWITCH_FIXED, A.CHOICE_PULL_CB, capacity=1 << 14)
    dg.close()


def test_cli_driver_replays_the_reference_sequence(tmp_path):
    """albba_bfs -m: read_Dmatrix -> convert_csr_to_csc -> run_custom_bfs_test
    through the reference entry points only; frontier sizes equal the
    reference's run_bfs_ref counts on the transposed matrix."""
    import subprocess

    exe = os.path.join(os.path.dirname(A.__file__), "albba_bfs")
    for name in ("general_real", "mixed_case_comments"):
        out = subprocess.run([exe, "-m", os.path.join(HERE, "golden", "mtx", name + ".mtx")],
                             capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        line = next(ln for ln in out.stdout.splitlines() if ln.startswith("frontier:"))
        got = np.array([int(x) for x in line.split()[1:]])
        np.testing.assert_array_equal(got, GG[f"dir_{name}__bfs_ref_frontier"])
    g = A.Graph.rmat(12, 16, seed=3, perm_seed=4)
    path = str(tmp_path / "g.cbd")
    g.write_cbd(path)
    roots = ",".join(str(r) for r in g.pick_roots(4, 1))
    out = subprocess.run([exe, "-m", path, "-r", roots], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.count("tree ok") == 4


@pytest.mark.parametrize("cap,blocks", [(0, 0), (1, 0), (8, 3), (64, 1), (1024, 7), (65536, 0),
                                        (65536, 1), (65536, 64)])
def test_small_frontier_push_runs(ctx, cap, blocks):
    """Runs of push levels inside the cooperative small-frontier kernel: edge
    caps that end a run after every level, after a few, or never, grids of 1
    to 64 workgroups, and the kernel switched off (cap 0) all give the
    oracle's levels, parents and per-level frontier sizes.  The RMAT graph
    switches to pull in the middle, so runs also start from a pulled
    frontier bitmap; repeated BFS on one device graph check that the kernel
    leaves its per-vertex keys reset."""
    for g in (A.Graph.grid(120, 90, 0.1, seed=51, perm_seed=52),
              A.Graph.rmat(13, 8, seed=53, perm_seed=54),
              A.Graph.er(3000, 3.0, seed=55, perm_seed=56)):
        rp, col = g.row_ptr.copy(), g.col.copy()
        dg = A.DeviceGraph(ctx, g.sellcs(64, 4096, 1))
        for root in g.pick_roots(3, 57, largest_cc=True):
            o = check(dg, rp, col, root, small_frontier_edges=cap, small_frontier_blocks=blocks,
                      capacity=1 << 12)
            assert (o["choice"][o["frontier"] > 0] >= 0).all()
            check(dg, rp, col, root, A.SWITCH_FIXED, A.CHOICE_PUSH, small_frontier_edges=cap)
        dg.close()


def test_bypass_counters(ctx):
    """Per-level bypass counts (reported, as the reference's bypass_slice and
    bypass_width): a pull skips the 64-row words that are off the unvisited
    list or fully visited, the column-bypass pull also skips the gathers into
    empty frontier words, and the plain pulls skip none; levels and parents
    stay those of the oracle."""
    g = A.Graph.rmat(15, 16, seed=121, perm_seed=122)
    rp, col = g.row_ptr.copy(), g.col.copy()
    dg = A.DeviceGraph(ctx, g.sellcs(64, 4096, 1))
    nwords = -(-g.n // 64)
    root = int(g.pick_roots(1, 123)[0])
    cb = check(dg, rp, col, root, A.SWITCH_FIXED, A.CHOICE_PULL_CB)
    sb = check(dg, rp, col, root, A.SWITCH_FIXED, A.CHOICE_PULL_SB)
    for o in (cb, sb):
        np.testing.assert_array_equal(o["bypassed"], nwords - o["words"])
        assert (o["bypassed"] >= 0).all() and o["bypassed"][-1] > 0
    assert (sb["col_bypassed"] == 0).all()
    assert (cb["col_bypassed"] <= cb["entries"]).all() and cb["col_bypassed"].sum() > 0
    # the same rows are scanned either way: a skipped gather never hits
    np.testing.assert_array_equal(cb["entries"], sb["entries"])
    push = check(dg, rp, col, root, A.SWITCH_FIXED, A.CHOICE_PUSH)
    assert (push["bypassed"] == 0).all() and (push["col_bypassed"] == 0).all()
    dg.close()


def test_bfs_multi_matches_single_runs(ctx):
    """albba_bfs_multi runs the roots back to back without a host
    synchronisation between them: every root's statistics equal those of a
    BFS of its own, the graph's results are the last root's (against the
    oracle), and per-root scripts are honoured."""
    g = A.Graph.rmat(15, 16, seed=131, perm_seed=132)
    rp, col = g.row_ptr.copy(), g.col.copy()
    dg = A.DeviceGraph(ctx, g.sellcs(64, 4096, 1))
    roots = [int(r) for r in g.pick_roots(7, 133)]
    rng = np.random.default_rng(134)
    scripts = [rng.integers(0, 4, 12).astype(np.int32) for _ in roots]
    try:
        for pol, scr in ((A.SWITCH_HEURISTIC, None), (A.SWITCH_SCRIPT, scripts)):
            many = dg.bfs_many(roots, policy=pol, scripts=scr)
            lv, pa = dg.results()
            wl, wp = O.bfs_queue(rp, col, roots[-1])
            np.testing.assert_array_equal(lv, wl)
            np.testing.assert_array_equal(pa, wp)
            for i, r in enumerate(roots):
                one = dg.bfs(r, policy=pol, script=None if scr is None else scr[i], results=False)
                for k in ("num_levels", "reached", "edges_in_component"):
                    assert many[k][i] == one[k], (k, i)
            assert (many["bfs_ms"] > 0).all()
        assert dg.bfs_many([])["bfs_ms"].size == 0
        with pytest.raises(A.AlbbaError) as e:
            dg.bfs_many([roots[0], g.n])
        assert e.value.code == -1
    finally:
        dg.close()


def test_swept_entries_of_list_pulls(ctx):
    """With per-level timing, a list pull reports the stored entries of the
    words it scanned (what a full scan of those slices reads): at least the
    entries it gathered, at most every stored entry, never growing from one
    pull level to the next, and at the root's level every entry but at most
    the root's own row; push levels report -1."""
    g = A.Graph.rmat(15, 16, seed=141, perm_seed=142)
    rp, col = g.row_ptr.copy(), g.col.copy()
    dg = A.DeviceGraph(ctx, g.sellcs(64, 4096, 1))
    root = int(g.pick_roots(1, 143)[0])
    deg = int(rp[root + 1] - rp[root])
    try:
        for ch in (A.CHOICE_PULL_SB, A.CHOICE_PULL_CB):
            o = check(dg, rp, col, root, A.SWITCH_FIXED, ch, time_levels=True)
            sw, en = o["swept"], o["entries"]
            assert sw[0] in (g.nnz, g.nnz - deg)
            assert (sw >= en).all() and (sw <= g.nnz).all()
            assert (np.diff(sw) <= 0).all()
        o = check(dg, rp, col, root, A.SWITCH_FIXED, A.CHOICE_PUSH, time_levels=True)
        assert (o["swept"] == -1).all()
        # without per-level timing nothing is counted
        o = dg.bfs(root, policy=A.SWITCH_FIXED, fixed_choice=A.CHOICE_PULL_CB, results=False)
        assert (o["swept"] == -1).all()
    finally:
        dg.close()
