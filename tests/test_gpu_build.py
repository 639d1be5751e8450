"""The SELL-C-sigma layout built on the device (albba_graph_build) against
the host builder (albba_sellcs_build, itself pinned to the reference
builder's golden arrays): every array bit for bit, for every slice height,
sigma and column padding, symmetric and directed graphs, rows out of order;
then BFS on the device-built graph against the oracle."""
import os

import numpy as np
import pytest

import albba_amd as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GG = np.load(os.path.join(HERE, "golden", "golden_graphs.npz"))
GOLDEN = sorted({k.split("__")[0] for k in GG.files})


def assert_same_layout(ctx, g, C_, sigma, pad, sort_first=False):
    host = g.sellcs(C_, sigma, pad, sort_first=sort_first).arrays()
    dg = A.DeviceGraph.build(ctx, g, C_, sigma, pad, sort_first=sort_first)
    dev = dg.layout().arrays()
    for k, v in host.items():
        if v is None:
            assert dev[k] is None, k
        else:
            np.testing.assert_array_equal(dev[k], v, err_msg=f"{k} C={C_} sigma={sigma} pad={pad}")
    return dg


@pytest.mark.parametrize("name", GOLDEN)
def test_golden_graphs_every_shape(ctx, name):
    rp, col = GG[f"{name}__rp"], GG[f"{name}__col"]
    g = A.Graph.from_arrays(rp, col, symmetric=not name.startswith("dir_"))
    for C_ in (16, 32, 64, 128):
        for sigma in sorted({C_, 256, 4096, g.n}):
            if sigma % C_ and sigma < g.n:
                continue
            for pad in (1, 16, 32):
                assert_same_layout(ctx, g, C_, sigma, pad).close()


@pytest.mark.parametrize("name", GOLDEN)
def test_golden_graphs_first_column_order(ctx, name):
    """The reference's -s 1 row order built on the device: the host
    builder's arrays, bit for bit; BFS on it (no sigma window holds its own
    id range, so level/parent go out through the position table) against
    the oracle; run_custom_bfs_test(sort_flag) gives the reference's
    frontier counts."""
    rp, col = GG[f"{name}__rp"], GG[f"{name}__col"]
    g = A.Graph.from_arrays(rp, col, symmetric=not name.startswith("dir_"))
    for C_, sigma, pad in [(64, 64, 1), (16, 32, 16), (128, 256, 1), (32, max(g.n, 32), 32)]:
        if sigma % C_ and sigma < g.n:
            continue
        dg = assert_same_layout(ctx, g, C_, sigma, pad, sort_first=True)
        for root in sorted({0, g.n // 2, g.n - 1}):
            o = dg.bfs(root)
            lv, pa = O.bfs_queue(rp, col, root)
            np.testing.assert_array_equal(o["level"], lv)
            np.testing.assert_array_equal(o["parent"], pa)
        dg.close()
    out = A.run_custom_bfs_test(rp.size - 1, rp.astype(np.int32), col, sort_flag=True)
    np.testing.assert_array_equal(out["frontier"], GG[f"{name}__bfs_ref_frontier"])


def test_first_column_order_medium_graph(ctx):
    g = A.Graph.rmat(15, 16, seed=101, perm_seed=102)
    for C_, sigma in [(64, 4096), (128, 16384)]:
        dg = assert_same_layout(ctx, g, C_, sigma, 1, sort_first=True)
        for root in g.pick_roots(2, 103):
            for pol, fc in ((A.SWITCH_HEURISTIC, 0), (A.SWITCH_FIXED, A.CHOICE_PULL_CB),
                            (A.SWITCH_FIXED, A.CHOICE_PUSH)):
                o = dg.bfs(int(root), policy=pol, fixed_choice=fc)
                lv, pa = O.bfs_queue(g.row_ptr, g.col, int(root))
                np.testing.assert_array_equal(o["level"], lv)
                np.testing.assert_array_equal(o["parent"], pa)
        dg.close()


@pytest.mark.parametrize("kind", ["rmat", "er", "grid", "directed"])
def test_medium_graphs_and_bfs(ctx, kind):
    if kind == "directed":
        base = A.Graph.rmat(13, 8, seed=91, perm_seed=92)
        rp, col = base.row_ptr, base.col
        # keep the entries u -> v with u < v of a symmetric graph: not symmetric
        rows = np.repeat(np.arange(base.n), np.diff(rp))
        keep = col > rows
        rp2 = np.zeros(base.n + 1, np.int64)
        np.cumsum(np.bincount(rows[keep], minlength=base.n), out=rp2[1:])
        g = A.Graph.from_arrays(rp2, col[keep].astype(np.int32), symmetric=False)
    else:
        g = {"rmat": lambda: A.Graph.rmat(14, 16, seed=93, perm_seed=94),
             "er": lambda: A.Graph.er(30000, 6.0, seed=95, perm_seed=96),
             "grid": lambda: A.Graph.grid(120, 130, 0.1, seed=97, perm_seed=98)}[kind]()
    for C_, sigma, pad in [(64, 4096, 1), (32, 256, 16), (128, g.n, 32), (16, 16, 1)]:
        dg = assert_same_layout(ctx, g, C_, sigma, pad)
        if (C_, sigma) == (64, 4096):
            assert dg.build_report["rows_sorted"]
            rp, col = g.row_ptr, g.col
            for root in g.pick_roots(3, 99):
                o = dg.bfs(int(root))
                lv, pa = O.bfs_queue(rp, col, int(root))
                np.testing.assert_array_equal(o["level"], lv)
                np.testing.assert_array_equal(o["parent"], pa)
        dg.close()


@pytest.mark.parametrize("directed", [False, True])
def test_position_tables_above_one_l2(ctx, directed):
    """Graphs of 1.1 M and 2.2 M vertices, whose position tables (16-bit
    in-window offsets with a power-of-two sigma, 32-bit positions otherwise)
    outgrow one 4 MB L2: layouts against the host builder, BFS against the
    oracle."""
    if directed:
        base = A.Graph.er(1_100_000, 4.0, seed=133, perm_seed=134)
        rp, col = base.row_ptr, base.col
        rows = np.repeat(np.arange(base.n), np.diff(rp))
        keep = col > rows
        rp2 = np.zeros(base.n + 1, np.int64)
        np.cumsum(np.bincount(rows[keep], minlength=base.n), out=rp2[1:])
        g = A.Graph.from_arrays(rp2, col[keep].astype(np.int32), symmetric=False)
        shapes = [(64, 960, 1), (32, 4096, 16)]
    else:
        g = A.Graph.er(2_200_000, 4.0, seed=131, perm_seed=132)
        shapes = [(64, 4096, 1), (128, 3072, 32)]
    for C_, sigma, pad in shapes:
        dg = assert_same_layout(ctx, g, C_, sigma, pad)
        root = int(g.pick_roots(1, 135)[0])
        o = dg.bfs(root)
        lv, pa = O.bfs_queue(g.row_ptr, g.col, root)
        np.testing.assert_array_equal(o["level"], lv)
        np.testing.assert_array_equal(o["parent"], pa)
        dg.close()


def test_rows_out_of_order_are_sorted(ctx):
    base = A.Graph.rmat(12, 16, seed=101, perm_seed=102)
    rp, col = base.row_ptr.copy(), base.col.copy()
    rng = np.random.default_rng(103)
    for v in range(base.n):
        rng.shuffle(col[rp[v]:rp[v + 1]])
    g = A.Graph.from_arrays(rp, col, symmetric=True)
    dg = assert_same_layout(ctx, g, 64, 4096, 1)
    assert not dg.build_report["rows_sorted"]
    root = int(base.pick_roots(1, 104)[0])
    o = dg.bfs(root)
    lv, pa = O.bfs_queue(base.row_ptr, base.col, root)
    np.testing.assert_array_equal(o["level"], lv)
    np.testing.assert_array_equal(o["parent"], pa)
    dg.close()


def test_empty_rows_and_tiny_graphs(ctx):
    # isolated vertices only; then one edge among many isolated vertices
    for rp, col in [(np.zeros(101, np.int64), np.zeros(0, np.int32)),
                    (np.r_[0, 1, 2, np.full(98, 2)].astype(np.int64), np.array([1, 0], np.int32))]:
        g = A.Graph.from_arrays(rp, col, symmetric=True)
        for C_, sigma, pad in [(64, 4096, 1), (16, 16, 32)]:
            assert_same_layout(ctx, g, C_, sigma, pad).close()


def star(n, directed):
    """Vertex 0 joined to every other vertex (rows of W list in-neighbours:
    directed, vertex 0's row is empty and every other row holds 0)."""
    if directed:
        rp = np.r_[0, np.arange(n)].astype(np.int64)
        col = np.zeros(n - 1, np.int32)
    else:
        rp = np.r_[0, n - 1, np.arange(n, 2 * n - 1)].astype(np.int64)
        col = np.r_[np.arange(1, n), np.zeros(n - 1)].astype(np.int32)
    return rp, col


@pytest.mark.parametrize("directed", [False, True])
def test_rows_longer_than_the_long_row_threshold(ctx, directed):
    # the hub's row (or, directed, its push row) holds 9999 entries
    rp, col = star(10000, directed)
    g = A.Graph.from_arrays(rp, col, symmetric=not directed)
    for C_, sigma, pad in [(64, 4096, 1), (128, 10000, 16), (16, 256, 32)]:
        dg = assert_same_layout(ctx, g, C_, sigma, pad)
        root = 0 if not directed else 5
        o = dg.bfs(root)
        lv, pa = O.bfs_queue(rp, col, root)
        np.testing.assert_array_equal(o["level"], lv)
        np.testing.assert_array_equal(o["parent"], pa)
        dg.close()


@pytest.mark.parametrize("directed", [False, True])
def test_heavy_row_groups(ctx, directed):
    """Every 64-position group holds more than 2 048 entries, so the row-major
    copy goes through the chunked pass for heavy groups; C = 64 takes the
    LDS-tiled column pass, the other heights the per-column one; directed,
    the push CSR comes from the sort-based transpose."""
    g = A.Graph.er(8192, 200.0, seed=71, perm_seed=72)
    rp, col = g.row_ptr, g.col
    if directed:
        # each row keeps the entries an asymmetric hash picks: about half
        rows = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(rp))
        keep = (rows * 2654435761 + col.astype(np.int64) * 40503) % 7 < 4
        rp = np.zeros(g.n + 1, np.int64)
        np.cumsum(np.bincount(rows[keep], minlength=g.n), out=rp[1:])
        col = g.col[keep].astype(np.int32)
        g = A.Graph.from_arrays(rp, col, symmetric=False)
    # the groups of C = 64, sigma = 4096: rows by length within each window
    lens = -np.sort(-np.diff(rp).reshape(-1, 4096), axis=1)
    assert lens.reshape(-1, 64).sum(axis=1).min() > 2048
    for C_, sigma, pad in [(64, 4096, 1), (64, 64, 16), (32, 256, 1), (128, g.n, 32)]:
        dg = assert_same_layout(ctx, g, C_, sigma, pad)
        if C_ == 64 and pad == 1:
            for root in g.pick_roots(2, 73):
                o = dg.bfs(int(root))
                lv, pa = O.bfs_queue(rp, col, int(root))
                np.testing.assert_array_equal(o["level"], lv)
                np.testing.assert_array_equal(o["parent"], pa)
        dg.close()


def test_build_errors(ctx):
    g = A.Graph.rmat(8, 4, seed=105, perm_seed=106)
    for C_, sigma, pad in [(48, 4096, 1), (64, 100, 1), (64, 4096, 7)]:
        with pytest.raises(A.AlbbaError):
            A.DeviceGraph.build(ctx, g, C_, sigma, pad)
    bad = g.col.copy()
    bad[5] = g.n  # a column index outside [0, n)
    with pytest.raises(A.AlbbaError, match="column index"):
        A.DeviceGraph.build(ctx, A.Graph.from_arrays(g.row_ptr, bad, symmetric=True), 64, 4096, 1)
    rp = g.row_ptr.copy()
    rp[3], rp[4] = rp[4], rp[3]  # decreasing row pointers
    if rp[3] != rp[4]:
        with pytest.raises(A.AlbbaError, match="row_ptr"):
            A.DeviceGraph.build(ctx, A.Graph.from_arrays(rp, g.col, symmetric=True), 64, 4096, 1)


def assert_same_as_host_path(ctx, path, C_, sigma, pad):
    host = A.Graph.read(str(path))
    ref = host.sellcs(C_, sigma, pad).arrays()
    dg = A.DeviceGraph.load(ctx, str(path), C_, sigma, pad)
    assert (dg.n, dg.nnz) == (host.n, host.nnz)
    got = dg.layout().arrays()
    for k, v in ref.items():
        if v is None:
            assert got[k] is None, k
        else:
            np.testing.assert_array_equal(got[k], v, err_msg=f"{k} C={C_} sigma={sigma} pad={pad}")
    return dg, host


@pytest.mark.parametrize("kind", ["rmat", "directed", "golden_dir"])
def test_cbd_load_equals_host_read_build_upload(ctx, tmp_path, kind):
    """.cbd straight into device memory, W = A^T formed and the layout built
    on the device: the same arrays as the host reader + builder + upload."""
    if kind == "rmat":
        g = A.Graph.rmat(12, 16, seed=111, perm_seed=112)
    elif kind == "directed":
        base = A.Graph.er(5000, 8.0, seed=113, perm_seed=114)
        rows = np.repeat(np.arange(base.n), np.diff(base.row_ptr))
        keep = (base.col * 7 + rows) % 3 == 0  # an arbitrary asymmetric subset
        rp = np.zeros(base.n + 1, np.int64)
        np.cumsum(np.bincount(rows[keep], minlength=base.n), out=rp[1:])
        g = A.Graph.from_arrays(rp, base.col[keep].astype(np.int32), symmetric=False)
    else:
        name = [k for k in GOLDEN if k.startswith("dir_")][0]
        g = A.Graph.from_arrays(GG[f"{name}__rp"], GG[f"{name}__col"], symmetric=False)
    path = tmp_path / "g.cbd"
    g.write_cbd(str(path))
    for C_, sigma, pad in [(64, 4096, 1), (32, 256, 16)]:
        dg, host = assert_same_as_host_path(ctx, path, C_, sigma, pad)
        for root in host.pick_roots(2, 115):
            o = dg.bfs(int(root))
            lv, pa = O.bfs_queue(host.row_ptr, host.col, int(root))
            np.testing.assert_array_equal(o["level"], lv)
            np.testing.assert_array_equal(o["parent"], pa)
        dg.close()


def test_cbd_load_errors(ctx, tmp_path):
    g = A.Graph.rmat(9, 8, seed=116, perm_seed=117)
    good = tmp_path / "g.cbd"
    g.write_cbd(str(good))
    raw = good.read_bytes()
    n, nnz = g.n, g.nnz
    # cut inside the column array
    cut = tmp_path / "cut.cbd"
    cut.write_bytes(raw[:12 + 4 * (n + 1) + 4 * (nnz // 2)])
    with pytest.raises(A.AlbbaError, match="short read"):
        A.DeviceGraph.load(ctx, str(cut))
    # a column index outside [0, n)
    bad = bytearray(raw)
    off = 12 + 4 * (n + 1) + 4 * 3
    bad[off:off + 4] = np.int32(n + 5).tobytes()
    badp = tmp_path / "bad.cbd"
    badp.write_bytes(bytes(bad))
    with pytest.raises(A.AlbbaError, match="out of range"):
        A.DeviceGraph.load(ctx, str(badp))
    # not square
    ns = bytearray(raw)
    ns[4:8] = np.int32(n + 1).tobytes()
    nsp = tmp_path / "ns.cbd"
    nsp.write_bytes(bytes(ns))
    with pytest.raises(A.AlbbaError):
        A.DeviceGraph.load(ctx, str(nsp))
    with pytest.raises(A.AlbbaError, match="not a .cbd"):
        A.DeviceGraph.load(ctx, str(tmp_path / "g.mtx"))
    # the good file still loads on the same context afterwards
    A.DeviceGraph.load(ctx, str(good)).close()


GEN_CASES = [
    ("rmat", dict(scale=10, edgefactor=16, seed=121, perm_seed=122), lambda: A.Graph.rmat(10, 16, seed=121, perm_seed=122)),
    ("rmat", dict(scale=12, edgefactor=8, a=0.45, b=0.25, c=0.15, seed=123, perm_seed=124),
     lambda: A.Graph.rmat(12, 8, 0.45, 0.25, 0.15, seed=123, perm_seed=124)),
    ("rmat", dict(scale=9, edgefactor=32, a=0.65, b=0.15, c=0.15, seed=125, perm_seed=126),
     lambda: A.Graph.rmat(9, 32, 0.65, 0.15, 0.15, seed=125, perm_seed=126)),
    ("er", dict(n=5000, mean_degree=6.0, seed=127, perm_seed=128), lambda: A.Graph.er(5000, 6.0, seed=127, perm_seed=128)),
    ("er", dict(n=3, mean_degree=40.0, seed=129, perm_seed=130), lambda: A.Graph.er(3, 40.0, seed=129, perm_seed=130)),
    ("grid", dict(rows=30, cols=41, p_remove=0.1, seed=131, perm_seed=132),
     lambda: A.Graph.grid(30, 41, 0.1, seed=131, perm_seed=132)),
    ("grid", dict(rows=7, cols=1, p_remove=0.5, seed=133, perm_seed=134),
     lambda: A.Graph.grid(7, 1, 0.5, seed=133, perm_seed=134)),
]


@pytest.mark.parametrize("case", range(len(GEN_CASES)))
def test_device_generators_equal_host_generators(ctx, case):
    """Edges, permutation, symmetrization, self-loop removal, sort and dedup
    on the device: the host generator's CSR bit for bit, then the layout."""
    kind, kw, host_fn = GEN_CASES[case]
    host = host_fn()
    dg, got = A.DeviceGraph.generate(ctx, kind, 64, 4096, 1, **kw)
    np.testing.assert_array_equal(got.row_ptr, host.row_ptr)
    np.testing.assert_array_equal(got.col, host.col)
    assert (dg.n, dg.nnz) == (host.n, host.nnz)
    ref = host.sellcs(64, 4096, 1).arrays()
    lay = dg.layout().arrays()
    for k, v in ref.items():
        if v is None:
            assert lay[k] is None
        else:
            np.testing.assert_array_equal(lay[k], v, err_msg=k)
    if host.nnz:
        root = int(host.pick_roots(1, 135)[0])
        o = dg.bfs(root)
        lv, pa = O.bfs_queue(host.row_ptr, host.col, root)
        np.testing.assert_array_equal(o["level"], lv)
        np.testing.assert_array_equal(o["parent"], pa)
    dg.close()


def test_device_generator_errors(ctx):
    for kind, kw in [("rmat", dict(scale=0)), ("rmat", dict(scale=8, a=0.6, b=0.3, c=0.3)),
                     ("er", dict(n=0, mean_degree=2.0)), ("grid", dict(rows=0, cols=4)),
                     ("grid", dict(rows=4, cols=4, p_remove=1.5))]:
        with pytest.raises(A.AlbbaError):
            A.DeviceGraph.generate(ctx, kind, **kw)
    with pytest.raises(A.AlbbaError):
        A.DeviceGraph.generate(ctx, "rmat", 48, 4096, 1, scale=8)
