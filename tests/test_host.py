"""Host-side product code (libalbba.so without a GPU): the biio loader, the
transpose, both SELL-C-sigma builders, the generators and the C ABI exports.
Expected values come from the reference fixtures and the pinned oracle."""
import ctypes as C
import os
import re
import shutil

import numpy as np
import pytest

import albba_amd as A
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
MTXDIR = os.path.join(HERE, "golden", "mtx")
GM = np.load(os.path.join(HERE, "golden", "golden_mtx.npz"))
GG = np.load(os.path.join(HERE, "golden", "golden_graphs.npz"))
MTX = sorted({k.split("__")[0] for k in GM.files})
GRAPHS = sorted({k.split("__")[0] for k in GG.files})
SELL_KEYS = ["row_order", "slice_widths", "vop_pointers", "vop_lengths", "slice_pointers",
             "column_indices", "max_index", "min_index"]


# ---- C ABI ---------------------------------------------------------------

def header_symbols():
    src = open(os.path.join(ROOT, "include", "albba.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    funcs = re.findall(r"ALBBA_API\s+[\w\s\*]+?\b(\w+)\s*\(", src)
    data = re.findall(r"extern\s+ALBBA_API\s+\w+\s+(\w+)\s*;", src)
    return sorted(set(funcs)), sorted(set(data))


def test_library_exports_every_header_symbol():
    funcs, data = header_symbols()
    assert len(funcs) > 40
    lib = A.lib()
    for name in funcs + data:
        assert hasattr(lib, name), f"{name} declared in include/albba.h but not exported"
    import subprocess

    nm = subprocess.run(["nm", "-D", "--defined-only", A.LIB_PATH], capture_output=True,
                        text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in nm.splitlines() if ln.strip()}
    for name in funcs + data:
        assert name in exported, name


def test_version_and_errors():
    assert A.lib().albba_version().startswith(b"albba")
    with pytest.raises(A.AlbbaError) as e:
        A.Graph.rmat(0)
    assert e.value.code == -1


# ---- loader -------------------------------------------------------------------

@pytest.mark.parametrize("name", MTX)
def test_read_Dmatrix_mtx_matches_reference(name):
    got = A.read_Dmatrix(os.path.join(MTXDIR, name + ".mtx"))
    nr, nc, nnz, sym = GM[f"{name}__shape"]
    assert got[:3] == (nr, nc, nnz)
    assert got[6] == sym
    np.testing.assert_array_equal(got[3], GM[f"{name}__row_ptr"])
    np.testing.assert_array_equal(got[4], GM[f"{name}__col_idx"])
    np.testing.assert_array_equal(got[5], GM[f"{name}__val"])


@pytest.mark.parametrize("name", MTX)
def test_convert_csr_to_csc_matches_reference(name):
    nr, nc, nnz, _ = GM[f"{name}__shape"]
    if nr != nc:
        pytest.skip("square only")
    cp, ri = A.convert_csr_to_csc(int(nr), GM[f"{name}__row_ptr"], GM[f"{name}__col_idx"])
    np.testing.assert_array_equal(cp, GM[f"{name}__csc_ptr"])
    np.testing.assert_array_equal(ri, GM[f"{name}__csc_idx"])


@pytest.mark.parametrize("name", MTX)
def test_read_graph_is_the_transposed_working_matrix(name):
    g = A.Graph.read(os.path.join(MTXDIR, name + ".mtx"))
    np.testing.assert_array_equal(g.row_ptr, GM[f"{name}__csc_ptr"])
    np.testing.assert_array_equal(g.col, GM[f"{name}__csc_idx"])


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
@pytest.mark.parametrize("name", MTX)
def test_read_Dmatrix_matches_live_reference(name):
    ref = O.ref_read(os.path.join(MTXDIR, name + ".mtx"))
    got = A.read_Dmatrix(os.path.join(MTXDIR, name + ".mtx"))
    assert ref[:3] == got[:3] and ref[6] == got[6]
    for a, b in zip(ref[3:6], got[3:6]):
        np.testing.assert_array_equal(a, b)


def test_cbd_round_trip_and_convert(tmp_path):
    g = A.Graph.rmat(9, 8, seed=5, perm_seed=6)
    p = str(tmp_path / "g.cbd")
    g.write_cbd(p)
    raw = open(p, "rb").read()
    hdr = np.frombuffer(raw[:12], np.int32)
    assert tuple(hdr) == (g.n, g.n, g.nnz)
    assert len(raw) == 12 + 4 * (g.n + 1) + 4 * g.nnz + 8 * g.nnz
    got = A.read_Dmatrix(p)
    np.testing.assert_array_equal(got[3], g.row_ptr)
    np.testing.assert_array_equal(got[4], g.col)
    assert np.all(got[5] == 1.0)
    # .mtx -> .cbd conversion (bfs/src/biio.h:98): the sibling .cbd appears
    src = tmp_path / "m.mtx"
    shutil.copy(os.path.join(MTXDIR, "symmetric_pattern.mtx"), src)
    L = A.lib()
    r, c_, z, s = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    rp, ci = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
    va = C.POINTER(C.c_double)()
    L.read_Dmatrix_convert(C.byref(r), C.byref(c_), C.byref(z), C.byref(rp), C.byref(ci),
                           C.byref(va), C.byref(s), str(src).encode())
    assert (tmp_path / "m.cbd").exists()
    again = A.read_Dmatrix(str(tmp_path / "m.cbd"))
    np.testing.assert_array_equal(again[3], GM["symmetric_pattern__row_ptr"])
    np.testing.assert_array_equal(again[4], GM["symmetric_pattern__col_idx"])


def test_loader_errors(tmp_path, capfd):
    assert A.read_Dmatrix(str(tmp_path / "missing.cbd")) is None
    assert A.read_Dmatrix(str(tmp_path / "missing.mtx")) is None
    assert A.read_Dmatrix(os.path.join(MTXDIR, "bad_banner.mtx")) is None
    (tmp_path / "x.txt").write_text("1 2 3\n")
    assert A.read_Dmatrix(str(tmp_path / "x.txt")) is None
    assert "unsupport" in capfd.readouterr().out
    short = tmp_path / "short.cbd"
    short.write_bytes(np.array([10, 10, 50], np.int32).tobytes() + b"\0" * 8)
    assert A.read_Dmatrix(str(short)) is None
    with pytest.raises(A.AlbbaError) as e:
        A.Graph.read(str(tmp_path / "missing.cbd"))
    assert e.value.code == -3
    with pytest.raises(A.AlbbaError):
        A.Graph.read(str(tmp_path / "x.txt"))


# ---- SELL-C-sigma ---------------------------------------------------------------

@pytest.mark.parametrize("name", GRAPHS)
@pytest.mark.parametrize("C_,sigma", [(256, 16384), (64, 4096), (64, 64), (32, 64),
                                      (128, 256), (16, 32)])
def test_sellcs_preprocess_matches_reference(name, C_, sigma):
    rp, col = GG[f"{name}__rp"], GG[f"{name}__col"]
    got = A.sellcs_preprocess(rp.size - 1, rp.astype(np.int32), col, C_, sigma)
    for k in SELL_KEYS:
        np.testing.assert_array_equal(got[k], GG[f"{name}__sell_{C_}_{sigma}__{k}"], err_msg=k)


@pytest.mark.parametrize("name", GRAPHS)
def test_sellcs_preprocess_sorted_matches_reference(name):
    rp, col = GG[f"{name}__rp"], GG[f"{name}__col"]
    got = A.sellcs_preprocess(rp.size - 1, rp.astype(np.int32), col, 256, 16384, sort_flag=True)
    for k in SELL_KEYS:
        np.testing.assert_array_equal(got[k], GG[f"{name}__sellsorted_256_16384__{k}"], err_msg=k)


@pytest.mark.parametrize("name", GRAPHS)
@pytest.mark.parametrize("C_,pad", [(64, 1), (64, 32), (16, 16), (32, 32), (128, 1), (128, 16)])
def test_device_layout_is_the_reference_layout_renumbered(name, C_, pad):
    """albba_sellcs_build = reference SELL (same C, sigma) with each column
    padded to `pad` entries (-1) and column ids mapped to positions."""
    rp, col = GG[f"{name}__rp"], GG[f"{name}__col"]
    g = A.Graph.from_arrays(rp, col, symmetric=not name.startswith("dir_"))
    n = rp.size - 1
    sigma = 64 if C_ <= 64 else 256
    ref = O.sell_build(rp, col, C_, sigma)
    s = g.sellcs(C_, sigma, pad).arrays()
    np.testing.assert_array_equal(s["row_order"], ref["row_order"])
    np.testing.assert_array_equal(s["pos_of"][s["row_order"]], np.arange(n))
    np.testing.assert_array_equal(s["slice_widths"], ref["slice_widths"])
    np.testing.assert_array_equal(s["vop_pointers"], ref["vop_pointers"])
    np.testing.assert_array_equal(s["vop_lengths"], ref["vop_lengths"])
    # unpad column by column
    ci = s["column_indices"]
    out, at = [], 0
    for act in s["vop_lengths"].astype(np.int64) + 1:
        padded = -(-act // pad) * pad
        assert np.all(ci[at + act:at + padded] == -1)
        out.append(ci[at:at + act])
        at += padded
    assert at == ci.size and at == s["slice_pointers"][-1]
    assert np.all(s["slice_pointers"] % pad == 0)
    flat = np.concatenate(out) if out else np.zeros(0, np.int32)
    np.testing.assert_array_equal(flat, s["pos_of"][ref["column_indices"]])
    deg = np.diff(rp)
    np.testing.assert_array_equal(s["row_lengths"][:n], deg[s["row_order"]])
    # push CSR = out-edges over positions
    pp, pc = s["push_ptr"], s["push_col"]
    pairs_push = {(int(u), int(v)) for u in range(len(pp) - 1) for v in pc[pp[u]:pp[u + 1]]}
    pairs_w = {(int(s["pos_of"][c]), int(s["pos_of"][v])) for v in range(n)
               for c in col[rp[v]:rp[v + 1]]}
    assert pairs_push == pairs_w
    # rows of the layout, row-major, in ascending vertex id: the pull CSR, or
    # the push CSR itself when the matrix is symmetric
    rptr, rcol = (s["pull_ptr"], s["pull_col"]) if s["pull_ptr"] is not None else (pp, pc)
    if name.startswith("dir_"):
        assert s["pull_ptr"] is not None
    for p in range(n):
        v = int(s["row_order"][p])
        np.testing.assert_array_equal(rcol[rptr[p]:rptr[p + 1]], s["pos_of"][col[rp[v]:rp[v + 1]]])


@pytest.mark.parametrize("name", GRAPHS)
@pytest.mark.parametrize("C_,sigma,pad", [(128, 16384, 1), (64, 64, 16), (16, 32, 1)])
def test_device_layout_first_column_order(name, C_, sigma, pad):
    """ALBBA_SELL_SORT_FIRST_COL (the reference's -s 1): the row order is the
    reference's sorted build (fixture at sigma 16384; the window sort does
    not depend on C), and the layout is the reference-compatible sorted
    builder's at the same C and sigma, renumbered to positions."""
    rp, col = GG[f"{name}__rp"], GG[f"{name}__col"]
    g = A.Graph.from_arrays(rp, col, symmetric=not name.startswith("dir_"))
    n = rp.size - 1
    s = g.sellcs(C_, sigma, pad, sort_first=True).arrays()
    assert s["order_flags"] == A.SELL_SORT_FIRST_COL
    ref = A.sellcs_preprocess(n, rp.astype(np.int32), col, C_, sigma, sort_flag=True)
    if sigma == 16384:
        np.testing.assert_array_equal(s["row_order"], GG[f"{name}__sellsorted_256_16384__row_order"])
    np.testing.assert_array_equal(s["row_order"], ref["row_order"])
    np.testing.assert_array_equal(s["pos_of"][s["row_order"]], np.arange(n))
    np.testing.assert_array_equal(s["slice_widths"], ref["slice_widths"])
    np.testing.assert_array_equal(s["vop_pointers"], ref["vop_pointers"])
    np.testing.assert_array_equal(s["vop_lengths"], ref["vop_lengths"])
    ci = s["column_indices"]
    out, at = [], 0
    for act in s["vop_lengths"].astype(np.int64) + 1:
        out.append(ci[at:at + act])
        at += -(-act // pad) * pad
    flat = np.concatenate(out) if out else np.zeros(0, np.int32)
    np.testing.assert_array_equal(flat, s["pos_of"][ref["column_indices"]])
    # without the flag the order differs on these graphs (or is the identity
    # on the ones whose first columns already descend)
    plain = g.sellcs(C_, sigma, pad).arrays()
    assert plain["order_flags"] == 0


def test_first_column_order_differs_from_plain_order():
    g = A.Graph.rmat(10, 8, seed=5, perm_seed=6)
    a = g.sellcs(64, 256, 1, sort_first=True).arrays()["row_order"]
    b = g.sellcs(64, 256, 1).arrays()["row_order"]
    assert not np.array_equal(a, b)
    # the window sort keeps each window's rows in first-column order on ties
    first = np.array([g.col[g.row_ptr[v]] if g.row_ptr[v] < g.nnz else 0 for v in range(g.n)])
    deg = np.diff(g.row_ptr)
    key = [(-deg[v], -first[v], v) for v in a[:256]]
    assert key == sorted(key)


def test_device_layout_rejects_bad_parameters():
    g = A.Graph.rmat(8, 8)
    for C_, sigma, pad in [(48, 96, 1), (64, 100, 1), (64, 4096, 8), (0, 64, 1)]:
        with pytest.raises(A.AlbbaError):
            g.sellcs(C_, sigma, pad)
    p = C.POINTER(A._SELL)()
    assert A.lib().albba_sellcs_build_ex(C.byref(g._raw), 64, 64, 1, 2, C.byref(p)) == -1


def test_unsorted_rows_are_sorted_for_the_layout():
    rp = np.array([0, 3, 4, 5, 6], np.int64)
    col = np.array([3, 1, 2, 0, 0, 0], np.int32)
    s = A.Graph.from_arrays(rp, col).sellcs(16, 16, 1).arrays()
    ro, po = s["row_order"], s["pos_of"]
    first = s["column_indices"][:s["vop_lengths"][0] + 1]
    row0_pos = int(np.nonzero(ro == 0)[0][0])
    assert row0_pos == 0
    # row 0's first stored neighbour is its smallest id (1)
    assert first[0] == po[1]


# ---- generators -------------------------------------------------------------------

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def np_rng(seed, idx):
    """splitmix64 finaliser of seed + (idx+1) * 0x9E3779B97F4A7C15 (wrapping)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.asarray(idx, np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def np_perm(n, seed):
    perm = np.arange(n)
    r = np_rng(seed, np.arange(n))
    for i in range(n - 1, 0, -1):
        j = int(r[i] % np.uint64(i + 1))
        perm[i], perm[j] = perm[j], perm[i]
    return perm


def np_sym_csr(n, src, dst):
    keep = src != dst
    u = np.concatenate([src[keep], dst[keep]])
    v = np.concatenate([dst[keep], src[keep]])
    key = np.unique(u.astype(np.int64) * n + v)
    u, v = key // n, key % n
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, u + 1, 1)
    return np.cumsum(rp), v.astype(np.int32)


def np_rmat(scale, ef, a, b, c, seed, pseed):
    n, ne = 1 << scale, ef << scale
    draws = (scale + 1) // 2
    e = np.arange(ne, dtype=np.uint64)
    u = np.zeros(ne, np.int64)
    v = np.zeros(ne, np.int64)
    for lv in range(scale):
        r = np_rng(seed, e * np.uint64(draws) + np.uint64(lv // 2))
        bits = (r & np.uint64(0xFFFFFFFF)) if lv & 1 else (r >> np.uint64(32))
        x = bits.astype(np.float64) / 4294967296.0
        bu = (x >= a + b).astype(np.int64)
        bv = (((x >= a) & (x < a + b)) | (x >= a + b + c)).astype(np.int64)
        u |= bu << (scale - 1 - lv)
        v |= bv << (scale - 1 - lv)
    p = np_perm(n, pseed)
    return np_sym_csr(n, p[u], p[v])


@pytest.mark.parametrize("scale,ef,a,b,c", [(8, 8, 0.57, 0.19, 0.19), (10, 16, 0.45, 0.25, 0.25),
                                            (9, 4, 0.65, 0.15, 0.15)])
def test_rmat_generator_follows_its_spec(scale, ef, a, b, c):
    g = A.Graph.rmat(scale, ef, a, b, c, seed=1234, perm_seed=99)
    rp, col = np_rmat(scale, ef, a, b, c, 1234, 99)
    np.testing.assert_array_equal(g.row_ptr, rp)
    np.testing.assert_array_equal(g.col, col)


def test_er_and_grid_generators_follow_their_spec():
    n, k = 700, 6.0
    g = A.Graph.er(n, k, seed=5, perm_seed=6)
    ne = int(n * k / 2 + 0.5)
    e = np.arange(ne, dtype=np.uint64)
    src = (np_rng(5, 2 * e) % np.uint64(n)).astype(np.int64)
    dst = (np_rng(5, 2 * e + np.uint64(1)) % np.uint64(n)).astype(np.int64)
    p = np_perm(n, 6)
    rp, col = np_sym_csr(n, p[src], p[dst])
    np.testing.assert_array_equal(g.row_ptr, rp)
    np.testing.assert_array_equal(g.col, col)

    R, Cc, pr = 13, 17, 0.2
    g = A.Graph.grid(R, Cc, pr, seed=4, perm_seed=2)
    H = R * (Cc - 1)
    e = np.arange(H + (R - 1) * Cc, dtype=np.uint64)
    keep = ((np_rng(4, e) >> np.uint64(11)).astype(np.float64) / 2.0 ** 53) >= pr
    ei = e.astype(np.int64)
    u = np.where(ei < H, (ei // (Cc - 1)) * Cc + ei % (Cc - 1), ei - H)
    v = np.where(ei < H, u + 1, u + Cc)
    p = np_perm(R * Cc, 2)
    rp, col = np_sym_csr(R * Cc, p[u[keep]], p[v[keep]])
    np.testing.assert_array_equal(g.row_ptr, rp)
    np.testing.assert_array_equal(g.col, col)


def test_generated_graphs_are_simple_symmetric_and_sorted():
    for g in [A.Graph.rmat(12, 16), A.Graph.er(3000, 8.0), A.Graph.grid(30, 40, 0.1)]:
        rp, col = g.row_ptr, g.col
        rows = np.repeat(np.arange(g.n), np.diff(rp))
        assert not np.any(rows == col)                      # no self loops
        for v in range(0, g.n, max(1, g.n // 200)):
            r = col[rp[v]:rp[v + 1]]
            assert np.all(np.diff(r) > 0)                    # sorted, deduplicated
        fw = set(zip(rows.tolist(), col.tolist()))
        assert all((b, a) in fw for a, b in list(fw)[:20000])  # symmetric
        assert g.symmetric


def test_rmat_degree_skew_grows_with_a():
    lo = np.sort(A.Graph.rmat(14, 16, 0.45, 0.25, 0.25).degrees())
    hi = np.sort(A.Graph.rmat(14, 16, 0.65, 0.15, 0.15).degrees())
    top = len(lo) // 100  # share of stored entries held by the top 1% of vertices
    assert hi[-top:].sum() / hi.sum() > 1.5 * lo[-top:].sum() / lo.sum()
    assert np.mean(hi == 0) > np.mean(lo == 0)


def test_pick_roots():
    g = A.Graph.rmat(12, 16)
    r1 = g.pick_roots(16, 3)
    assert len(r1) == 16 and len(set(r1.tolist())) == 16
    assert np.all(g.degrees()[r1] > 0)
    np.testing.assert_array_equal(r1, g.pick_roots(16, 3))
    assert not np.array_equal(r1, g.pick_roots(16, 4))
    # largest component only
    rp, col = GG["two_comp__rp"], GG["two_comp__col"]
    tc = A.Graph.from_arrays(rp, col, symmetric=True)
    lv, _ = O.bfs_queue(rp, col, 0)
    big = set(np.nonzero(lv >= 0)[0].tolist())
    roots = tc.pick_roots(5, 1, largest_cc=True)
    sizes = [int((O.bfs_queue(rp, col, int(r))[0] >= 0).sum()) for r in roots]
    assert len(set(sizes)) == 1
    assert tc.pick_roots(1000, 1).size == int((tc.degrees() > 0).sum())
    assert big  # component of vertex 0 exists


def test_selector_queue_threshold():
    """The trainer's queue-listing ratio: the out-edges-per-vertex threshold
    that misplaces the fewest push levels (followed by a push above it, or by
    a pull at or below it), ties to the larger threshold."""
    from albba_amd.train_selector import queue_threshold

    assert queue_threshold([]) == 0.0
    sep = [(0.001, True), (0.002, True), (0.01, False), (0.02, False)]
    assert queue_threshold(sep) == 0.002
    # one misplaced level either way at 0 and at 0.002: the larger wins
    assert queue_threshold([(0.001, False), (0.002, True)]) == 0.002
    # every push followed by a pull: nothing needs listing
    assert queue_threshold([(0.003, False), (0.004, False)]) == 0.0
    # every push followed by a push: list up to the largest
    assert queue_threshold([(0.003, True), (0.004, True)]) == 0.004
