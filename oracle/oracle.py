"""ctypes binding of the CPU oracle (liboracle.so) and of the reference build
(_ref/libalbba_ref.so).  TEST INFRASTRUCTURE ONLY: imported by tests/, by the
smoke check and by bench.py's CPU-baseline leg, never by the product."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libalbba_ref.so")

I32P = C.POINTER(C.c_int32)
I64P = C.POINTER(C.c_int64)
U8P = C.POINTER(C.c_uint8)


class OrcSell(C.Structure):
    _fields_ = [("n", C.c_int32), ("C", C.c_int32), ("sigma", C.c_int32), ("nslices", C.c_int32),
                ("nvops", C.c_int32), ("nnz", C.c_int64), ("row_order", I32P),
                ("slice_widths", I32P), ("vop_pointers", I32P), ("vop_lengths", U8P),
                ("slice_pointers", I32P), ("column_indices", I32P), ("max_index", I32P),
                ("min_index", I32P)]


_orc = None
_ref = None


def build(ref: bool = False) -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def orc() -> C.CDLL:
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        L.orc_sell_build.argtypes = [C.c_int32, I64P, I32P, C.c_int32, C.c_int32, C.POINTER(OrcSell)]
        L.orc_sell_build.restype = C.c_int
        L.orc_sell_free.argtypes = [C.POINTER(OrcSell)]
        L.orc_sell_free.restype = None
        L.orc_bfs_queue.argtypes = [C.c_int32, I64P, I32P, C.c_int32, I32P, I32P]
        L.orc_bfs_queue.restype = C.c_int
        L.orc_bfs_dense_ref.argtypes = [C.c_int32, I64P, I32P, C.c_int32, I32P, I32P]
        L.orc_bfs_dense_ref.restype = C.c_int
        L.orc_albba_bfs.argtypes = [C.POINTER(OrcSell), C.c_int32, C.c_int32, I32P, I32P, I32P, I32P,
                                    I64P, I64P, C.c_int32]
        L.orc_albba_bfs.restype = C.c_int
        L.orc_albba_bfs_levels.argtypes = [C.POINTER(OrcSell), C.c_int32, C.c_int32, I32P, I32P, I32P,
                                           I32P, I64P, I64P, C.c_int32, C.c_int32]
        L.orc_albba_bfs_levels.restype = C.c_int
        L.orc_albba_bfs_f9.argtypes = [C.POINTER(OrcSell), C.c_int32, C.c_int32, I32P, I32P, I32P, I32P,
                                       C.c_int32]
        L.orc_albba_bfs_f9.restype = C.c_int
        L.orc_validate.argtypes = [C.c_int32, I64P, I32P, C.c_int32, I32P, I32P]
        L.orc_validate.restype = C.c_int
        _orc = L
    return _orc


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        L = C.CDLL(REF_SO)
        PI = C.POINTER(C.c_int)
        PPI = C.POINTER(PI)
        L.refh_read.argtypes = [C.c_char_p, PI, PI, PI, PPI, PPI, C.POINTER(C.POINTER(C.c_double)), PI]
        L.refh_read.restype = C.c_int
        L.refh_free.argtypes = [C.c_void_p]
        L.refh_free.restype = None
        L.refh_csc.argtypes = [C.c_int, C.c_int, C.c_int, PI, PI, PPI, PPI]
        L.refh_csc.restype = None
        L.refh_sell.argtypes = [C.c_int, C.c_int, PI, PI, C.c_int, C.c_int, C.c_int, PI, PPI, PPI,
                                PPI, C.POINTER(C.POINTER(C.c_ubyte)), PPI, PPI, PPI, PPI]
        L.refh_sell.restype = C.c_int
        L.refh_bfs_ref.argtypes = [C.c_int, C.c_int, PI, PI, PI, PI, PI]
        L.refh_bfs_ref.restype = C.c_int
        _ref = L
    return _ref


def _arr(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _take(ptr, n, dt):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dt)
    return np.ctypeslib.as_array(ptr, shape=(int(n),)).view(dt).copy()


# ---- oracle ----------------------------------------------------------------


def sell_build(row_ptr, col, C_: int, sigma: int) -> dict:
    rp, cc = _arr(row_ptr, np.int64), _arr(col, np.int32)
    n = rp.size - 1
    s = OrcSell()
    rc = orc().orc_sell_build(n, rp.ctypes.data_as(I64P), cc.ctypes.data_as(I32P), C_, sigma, C.byref(s))
    if rc:
        raise RuntimeError(f"orc_sell_build failed: {rc}")
    out = {
        "nslices": s.nslices,
        "row_order": _take(s.row_order, n, np.int32),
        "slice_widths": _take(s.slice_widths, s.nslices, np.int32),
        "vop_pointers": _take(s.vop_pointers, s.nslices + 1, np.int32),
        "vop_lengths": _take(s.vop_lengths, s.nvops, np.uint8),
        "slice_pointers": _take(s.slice_pointers, s.nslices + 1, np.int32),
        "column_indices": _take(s.column_indices, s.nnz, np.int32),
        "max_index": _take(s.max_index, s.nvops, np.int32),
        "min_index": _take(s.min_index, s.nvops, np.int32),
    }
    orc().orc_sell_free(C.byref(s))
    return out


def bfs_queue(row_ptr, col, root: int):
    rp, cc = _arr(row_ptr, np.int64), _arr(col, np.int32)
    n = rp.size - 1
    lv = np.empty(n, np.int32)
    pa = np.empty(n, np.int32)
    k = orc().orc_bfs_queue(n, rp.ctypes.data_as(I64P), cc.ctypes.data_as(I32P), root,
                            lv.ctypes.data_as(I32P), pa.ctypes.data_as(I32P))
    if k < 0:
        raise RuntimeError(f"orc_bfs_queue failed: {k}")
    return lv, pa


def bfs_dense_ref(row_ptr, col, root: int):
    rp, cc = _arr(row_ptr, np.int64), _arr(col, np.int32)
    n = rp.size - 1
    fr = np.zeros(n + 2, np.int32)
    lv = np.empty(n, np.int32)
    k = orc().orc_bfs_dense_ref(n, rp.ctypes.data_as(I64P), cc.ctypes.data_as(I32P), root,
                                fr.ctypes.data_as(I32P), lv.ctypes.data_as(I32P))
    if k < 0:
        raise RuntimeError(f"orc_bfs_dense_ref failed: {k}")
    return fr[:k].copy(), lv


def albba_bfs(row_ptr, col, root: int, C_: int = 256, sigma: int = 16384, tasks: int = 8,
              capacity: int = 1 << 16):
    """ALBBA level loop restated on the CPU (OpenMP over `tasks`)."""
    rp, cc = _arr(row_ptr, np.int64), _arr(col, np.int32)
    n = rp.size - 1
    s = OrcSell()
    rc = orc().orc_sell_build(n, rp.ctypes.data_as(I64P), cc.ctypes.data_as(I32P), C_, sigma, C.byref(s))
    if rc:
        raise RuntimeError(f"orc_sell_build failed: {rc}")
    try:
        return albba_bfs_on(s, n, root, tasks, capacity)
    finally:
        orc().orc_sell_free(C.byref(s))


def albba_bfs_on(s: OrcSell, n: int, root: int, tasks: int = 8, capacity: int = 1 << 16,
                 max_levels: int = 2 ** 31 - 1):
    """max_levels: stop after that many levels (a bounded timing sample)."""
    lv = np.empty(n, np.int32)
    pa = np.empty(n, np.int32)
    fr = np.zeros(capacity, np.int32)
    md = np.zeros(capacity, np.int32)
    bs = np.zeros(capacity, np.int64)
    bc = np.zeros(capacity, np.int64)
    k = orc().orc_albba_bfs_levels(C.byref(s), root, tasks, lv.ctypes.data_as(I32P),
                                   pa.ctypes.data_as(I32P), fr.ctypes.data_as(I32P),
                                   md.ctypes.data_as(I32P), bs.ctypes.data_as(I64P),
                                   bc.ctypes.data_as(I64P), capacity, max_levels)
    if k < 0:
        raise RuntimeError(f"orc_albba_bfs failed: {k}")
    m = min(k, capacity)
    return {"level": lv, "parent": pa, "num_levels": k, "frontier": fr[:m].copy(),
            "mode": md[:m].copy(), "bypass_slice": bs[:m].copy(), "bypass_col": bc[:m].copy()}


def albba_bfs_f9(s: OrcSell, n: int, root: int, tasks: int = 8, capacity: int = 1 << 16):
    """The level loop with the reference's dense2sparse offsets (F9): shows
    the defect, never used as a reference result."""
    lv = np.empty(n, np.int32)
    pa = np.empty(n, np.int32)
    fr = np.zeros(capacity, np.int32)
    md = np.zeros(capacity, np.int32)
    k = orc().orc_albba_bfs_f9(C.byref(s), root, tasks, lv.ctypes.data_as(I32P), pa.ctypes.data_as(I32P),
                               fr.ctypes.data_as(I32P), md.ctypes.data_as(I32P), capacity)
    if k < 0:
        raise RuntimeError(f"orc_albba_bfs_f9 failed: {k}")
    m = min(k, capacity)
    return {"level": lv, "parent": pa, "num_levels": k, "frontier": fr[:m].copy(), "mode": md[:m].copy()}


def build_sell(row_ptr, col, C_: int, sigma: int):
    """Keep an oracle SELL alive for repeated albba_bfs_on calls."""
    rp, cc = _arr(row_ptr, np.int64), _arr(col, np.int32)
    s = OrcSell()
    rc = orc().orc_sell_build(rp.size - 1, rp.ctypes.data_as(I64P), cc.ctypes.data_as(I32P), C_,
                              sigma, C.byref(s))
    if rc:
        raise RuntimeError(f"orc_sell_build failed: {rc}")
    return s


def free_sell(s: OrcSell) -> None:
    orc().orc_sell_free(C.byref(s))


def validate(row_ptr, col, root: int, level, parent) -> int:
    rp, cc = _arr(row_ptr, np.int64), _arr(col, np.int32)
    lv, pa = _arr(level, np.int32), _arr(parent, np.int32)
    return int(orc().orc_validate(rp.size - 1, rp.ctypes.data_as(I64P), cc.ctypes.data_as(I32P), root,
                                  lv.ctypes.data_as(I32P), pa.ctypes.data_as(I32P)))


# ---- reference build ---------------------------------------------------------


def ref_read(path: str):
    L = ref()
    r, c, z, s = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    rp, ci = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
    va = C.POINTER(C.c_double)()
    rc = L.refh_read(path.encode(), C.byref(r), C.byref(c), C.byref(z), C.byref(rp), C.byref(ci),
                     C.byref(va), C.byref(s))
    if rc:
        return None
    out = (r.value, c.value, z.value, _take(rp, r.value + 1, np.int32), _take(ci, z.value, np.int32),
           _take(va, z.value, np.float64), s.value)
    for p in (rp, ci, va):
        L.refh_free(C.cast(p, C.c_void_p))
    return out


def ref_csc(nr: int, nc: int, row_ptr, col):
    L = ref()
    rp, cc = _arr(row_ptr, np.int32), _arr(col, np.int32)
    cp, ri = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
    L.refh_csc(nr, nc, cc.size, rp.ctypes.data_as(C.POINTER(C.c_int)),
               cc.ctypes.data_as(C.POINTER(C.c_int)), C.byref(cp), C.byref(ri))
    out = (_take(cp, nc + 1, np.int32), _take(ri, cc.size, np.int32))
    L.refh_free(C.cast(cp, C.c_void_p))
    L.refh_free(C.cast(ri, C.c_void_p))
    return out


def ref_sell(row_ptr, col, C_: int, sigma: int, sort_flag: bool = False) -> dict:
    L = ref()
    rp, cc = _arr(row_ptr, np.int32), _arr(col, np.int32)
    n = rp.size - 1
    PI = C.POINTER(C.c_int)
    ns = C.c_int()
    ro, wd, vp, sp, ci, mx, mn = (PI() for _ in range(7))
    vl = C.POINTER(C.c_ubyte)()
    L.refh_sell(n, cc.size, rp.ctypes.data_as(PI), cc.ctypes.data_as(PI), C_, sigma, int(sort_flag),
                C.byref(ns), C.byref(ro), C.byref(wd), C.byref(vp), C.byref(vl), C.byref(sp),
                C.byref(ci), C.byref(mx), C.byref(mn))
    nsl = ns.value
    vpa = _take(vp, nsl + 1, np.int32)
    nv = int(vpa[-1])
    out = {
        "nslices": nsl,
        "row_order": _take(ro, n, np.int32),
        "slice_widths": _take(wd, nsl, np.int32),
        "vop_pointers": vpa,
        "vop_lengths": _take(vl, nv, np.uint8),
        "slice_pointers": _take(sp, nsl + 1, np.int32),
        "column_indices": _take(ci, cc.size, np.int32),
        "max_index": _take(mx, nv, np.int32),
        "min_index": _take(mn, nv, np.int32),
    }
    for p in (ro, wd, vp, vl, sp, ci, mx, mn):
        L.refh_free(C.cast(p, C.c_void_p))
    return out


def ref_bfs_frontiers(row_ptr, col):
    """run_bfs_ref (bfs/src/bfs_ref.h:26) from vertex 0: num_vex_ref[:iters]."""
    L = ref()
    rp, cc = _arr(row_ptr, np.int32), _arr(col, np.int32)
    n = rp.size - 1
    fr = np.zeros(n + 2, np.int32)
    it, sv = C.c_int(), C.c_int()
    PI = C.POINTER(C.c_int)
    L.refh_bfs_ref(n, cc.size, rp.ctypes.data_as(PI), cc.ctypes.data_as(PI), fr.ctypes.data_as(PI),
                   C.byref(it), C.byref(sv))
    return fr[:it.value].copy(), sv.value
