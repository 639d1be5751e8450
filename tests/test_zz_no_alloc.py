"""albba_batch_run_noalloc (the batched BFS without any allocation) through
the C ABI against albba_batch_run: every instance's levels and min-id
parents bit for bit and the per-instance results equal, on two calls in a
row (the second on a smaller batch), for both direction policies, from the
same team split (each call after albba_batch_forget) and again with the
teams sized from the previous run; batches beyond ALBBA_BATCH_NOALLOC_MAX
instances are rejected.  The oracle's queue BFS checks the no-alloc results
as well."""
import ctypes as C
import os

import numpy as np
import pytest

import albba_amd as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu
NOALLOC_MAX = 1024  # ALBBA_BATCH_NOALLOC_MAX
RESULT_FIELDS = ("num_levels", "error", "reached", "edges_in_component", "push_levels", "pull_levels")
FIELDS = RESULT_FIELDS + ("team_size", "team")


def run_abi(b, roots, noalloc, policy=A.SWITCH_HEURISTIC, selector=None, mapping=None, forget=True):
    """One call of albba_batch_run or albba_batch_run_noalloc through ctypes;
    with `forget`, from the entries-share team split (the previous run's
    levels and times dropped first), so that both entry points split alike."""
    L = A.lib()
    if forget:
        L.albba_batch_forget(b._p)
    n = len(b.graphs)
    r = np.ascontiguousarray(roots, dtype=np.int32)
    o = A._Opts()
    L.albba_bfs_opts_default(C.byref(o))
    o.switch_policy = policy
    if selector is not None:
        o.selector = selector._p
    mp = None if mapping is None else np.ascontiguousarray(mapping, dtype=np.int32)
    res = (A._BatchResult * n)()
    ms = C.c_float()
    fn = L.albba_batch_run_noalloc if noalloc else L.albba_batch_run
    rc = fn(b._p, r.ctypes.data_as(A.I32P), C.byref(o), mp.ctypes.data_as(A.I32P) if mp is not None else None,
            0, res, C.byref(ms))
    out = {k: np.array([getattr(x, k) for x in res]) for k in FIELDS}
    out["rc"] = rc
    out["levels"] = [b.results(i) for i in range(n)]
    return out


def same(a, b, fields=FIELDS):
    assert a["rc"] == 0 and b["rc"] == 0
    for k in fields:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    for i, ((la, pa), (lb, pb)) in enumerate(zip(a["levels"], b["levels"])):
        np.testing.assert_array_equal(la, lb, err_msg=f"instance {i}")
        np.testing.assert_array_equal(pa, pb, err_msg=f"instance {i}")


def test_noalloc_batch_matches_allocating_batch(ctx):
    big = [A.Graph.rmat(13, 16, seed=201, perm_seed=202), A.Graph.er(30000, 8.0, seed=203, perm_seed=204),
           A.Graph.grid(200, 150, 0.1, seed=205, perm_seed=206), A.Graph.rmat(11, 8, seed=207, perm_seed=208)]
    small = [A.Graph.rmat(10, 8, seed=209, perm_seed=210), A.Graph.grid(40, 30, 0.1, seed=211, perm_seed=212)]
    sel = A.Selector(ctx)
    sel.load(os.path.join(os.path.dirname(A.__file__), "selector.txt"))
    try:
        for graphs, copies in ((big, 5), (small, 2)):
            dgs = [A.DeviceGraph(ctx, g.sellcs(64, 4096, 1)) for g in graphs]
            inst = [k % len(graphs) for k in range(len(graphs) * copies)]
            roots = [int(graphs[gi].pick_roots(1, 300 + k, largest_cc=True)[0]) for k, gi in enumerate(inst)]
            b = A.Batch(ctx, [dgs[gi] for gi in inst])
            try:
                for pol, s in ((A.SWITCH_HEURISTIC, None), (A.SWITCH_LEARNED, sel)):
                    ref = run_abi(b, roots, False, pol, s)
                    got = run_abi(b, roots, True, pol, s)
                    same(ref, got)
                    # teams sized from the run before, still no allocation
                    hist = run_abi(b, roots, True, pol, s, forget=False)
                    same(ref, hist, RESULT_FIELDS)
                    for k, gi in enumerate(inst):
                        lv, pa = O.bfs_queue(graphs[gi].row_ptr, graphs[gi].col, roots[k])
                        np.testing.assert_array_equal(got["levels"][k][0], lv)
                        np.testing.assert_array_equal(got["levels"][k][1], pa)
            finally:
                b.close()
                for dg in dgs:
                    dg.close()
    finally:
        sel.close()


def test_noalloc_batch_rejects_more_than_its_maximum(ctx):
    g = A.Graph.grid(8, 8, 0.0, seed=1, perm_seed=2)
    dg = A.DeviceGraph(ctx, g.sellcs(16, 16, 1))
    b = A.Batch(ctx, [dg] * (NOALLOC_MAX + 1))
    try:
        roots = np.zeros(NOALLOC_MAX + 1, np.int32)
        out = run_abi(b, roots, True)
        assert out["rc"] == -6  # ALBBA_ERANGE
        assert b"at most 1024 instances" in A.lib().albba_last_error()
        # the allocating form runs the same batch
        assert run_abi(b, roots, False)["rc"] == 0
    finally:
        b.close()
        dg.close()
