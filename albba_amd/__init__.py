"""Python binding of libalbba.so, the MI355X-native ALBBA linear-algebra BFS.

The product is the C ABI declared in ``include/albba.h``; this module only
binds it with ctypes for tests, benchmarks and Python callers.  It mirrors the
reference's entry points on the graph-load -> SELL-C-sigma -> BFS path
(``read_Dmatrix`` bfs/src/biio.h:127, ``convert_csr_to_csc``
bfs/src/sparse_matrix.c:230, ``sellcs_preprocess`` bfs/src/bfs_SELLCSv2.c:34,
``run_custom_bfs_test`` bfs/src/bfs.h:13) and adds the root-aware GPU BFS,
the generators, the learned selector and the batch/autotune helpers.

Nothing here falls back to a CPU implementation: every BFS runs the HIP
kernels of libalbba.so, and a missing library or GPU raises.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libalbba.so")

CHOICE_PUSH, CHOICE_PULL, CHOICE_PULL_SB, CHOICE_PULL_CB = 0, 1, 2, 3
CHOICE_NAMES = ("push", "pull", "pull_sb", "pull_cb")
SWITCH_HEURISTIC, SWITCH_LEARNED, SWITCH_FIXED, SWITCH_SCRIPT, SWITCH_BEST = 0, 1, 2, 3, 4
GEN_RMAT, GEN_ER, GEN_GRID = 0, 1, 2
SELL_SORT_FIRST_COL = 1  # ALBBA_SELL_SORT_FIRST_COL: the reference's -s 1 row order
NUM_FEATURES = 12
MLP_NPARAM = 12 * 32 + 32 + 32 * 32 + 32 + 32 * 4 + 4
ERRORS = {
    -1: "EINVAL", -2: "ENOMEM", -3: "EIO", -4: "EHIP", -5: "ENODEV", -6: "ERANGE", -7: "EDEVMEM",
}

I32P = C.POINTER(C.c_int32)
I64P = C.POINTER(C.c_int64)
U8P = C.POINTER(C.c_uint8)
F32P = C.POINTER(C.c_float)
F64P = C.POINTER(C.c_double)


class AlbbaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"albba error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class _CSR(C.Structure):
    _fields_ = [("n", C.c_int32), ("nnz", C.c_int64), ("row_ptr", I64P), ("col", I32P),
                ("symmetric", C.c_int32)]


class _SELL(C.Structure):
    _fields_ = [("n", C.c_int32), ("C", C.c_int32), ("sigma", C.c_int32), ("pad", C.c_int32),
                ("nslices", C.c_int32), ("nnz", C.c_int64), ("padded_nnz", C.c_int64),
                ("nvops", C.c_int64), ("row_order", I32P), ("pos_of", I32P),
                ("slice_widths", I32P), ("slice_pointers", I64P), ("vop_pointers", I64P),
                ("vop_lengths", U8P), ("column_indices", I32P), ("row_lengths", I32P),
                ("push_ptr", I64P), ("push_col", I32P), ("pull_ptr", I64P), ("pull_col", I32P),
                ("order_flags", C.c_int32)]


class SparseMatrixCSR(C.Structure):
    """bfs/src/sparse_matrix.h:18-28."""
    _fields_ = [("name", C.c_char_p), ("values", F64P), ("column_indices", C.POINTER(C.c_int)),
                ("row_pointers", C.POINTER(C.c_int)), ("nrows", C.c_int), ("ncolumns", C.c_int),
                ("nnz", C.c_int), ("isSymmetric", C.c_int)]


class SparseMatrixCSC(C.Structure):
    _fields_ = [("name", C.c_char_p), ("values", F64P), ("row_indices", C.POINTER(C.c_int)),
                ("column_pointers", C.POINTER(C.c_int)), ("nrows", C.c_int), ("ncolumns", C.c_int),
                ("nnz", C.c_int)]


class SparseMatrixSELLCS(C.Structure):
    """bfs/src/sparse_matrix.h:43-63."""
    _fields_ = [("name", C.c_char_p), ("values", F64P), ("column_indices", C.POINTER(C.c_int)),
                ("nrows", C.c_int), ("ncolumns", C.c_int), ("nnz", C.c_int), ("C", C.c_int),
                ("sigma", C.c_int), ("nslices", C.c_int), ("slice_widths", C.POINTER(C.c_int)),
                ("slice_pointers", C.POINTER(C.c_int)), ("row_order", C.POINTER(C.c_int)),
                ("vop_lengths", U8P), ("vop_pointers", C.POINTER(C.c_int)),
                ("slice_widths_pointers", C.POINTER(C.c_int)), ("max_index", C.POINTER(C.c_int)),
                ("min_index", C.POINTER(C.c_int))]


class _Opts(C.Structure):
    _fields_ = [("switch_policy", C.c_int32), ("fixed_choice", C.c_int32), ("script", I32P),
                ("script_len", C.c_int32), ("selector", C.c_void_p), ("time_levels", C.c_int32),
                ("timing_reps", C.c_int32), ("small_frontier_edges", C.c_int64),
                ("small_frontier_blocks", C.c_int32), ("reserved_", C.c_int32)]


class _Stats(C.Structure):
    _fields_ = [("num_levels", C.c_int32), ("capacity", C.c_int32), ("reached", C.c_int64),
                ("edges_in_component", C.c_int64), ("bfs_ms", C.c_double), ("frontier", I32P),
                ("choice", I32P), ("level_ms", F32P), ("features", F32P), ("choice_ms", F32P),
                ("entries", I64P), ("words", I64P), ("bypassed", I64P), ("kernel_ms", F32P),
                ("col_bypassed", I64P), ("swept", I64P)]


class _BatchResult(C.Structure):
    _fields_ = [("num_levels", C.c_int32), ("error", C.c_int32), ("reached", C.c_int64),
                ("edges_in_component", C.c_int64), ("ms", C.c_double), ("done_ms", C.c_double),
                ("push_levels", C.c_int32), ("pull_levels", C.c_int32), ("team_size", C.c_int32),
                ("team", C.c_int32)]


class _BuildReport(C.Structure):
    _fields_ = [("upload_ms", C.c_float), ("build_ms", C.c_float), ("rows_sorted", C.c_int32),
                ("reserved_", C.c_int32)]


class _GenParams(C.Structure):
    _fields_ = [("kind", C.c_int32), ("scale", C.c_int32), ("edgefactor", C.c_int32),
                ("a", C.c_double), ("b", C.c_double), ("c", C.c_double), ("n", C.c_int32),
                ("rows", C.c_int32), ("cols", C.c_int32), ("mean_degree", C.c_double),
                ("p_remove", C.c_double), ("seed", C.c_uint64), ("perm_seed", C.c_uint64)]


class _GenReport(C.Structure):
    _fields_ = [("perm_ms", C.c_float), ("gen_ms", C.c_float), ("build_ms", C.c_float),
                ("total_ms", C.c_float)]


class _LoadReport(C.Structure):
    _fields_ = [("read_ms", C.c_float), ("prep_ms", C.c_float), ("build_ms", C.c_float),
                ("total_ms", C.c_float)]


class _TuneConfig(C.Structure):
    _fields_ = [("C", C.c_int32), ("sigma", C.c_int32), ("mapping", C.c_int32),
                ("status", C.c_int32), ("ms", C.c_double), ("batch_ms", C.c_double)]


_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    """Load libalbba.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(make -C albba_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    sig = {
        "albba_last_error": (C.c_char_p, []),
        "albba_version": (C.c_char_p, []),
        "albba_rng_u64": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "albba_csr_free": (None, [C.POINTER(_CSR)]),
        "albba_read_graph": (C.c_int, [C.c_char_p, C.POINTER(_CSR)]),
        "albba_write_cbd": (C.c_int, [C.c_char_p, C.POINTER(_CSR)]),
        "albba_gen_rmat": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                     C.c_uint64, C.c_uint64, C.POINTER(_CSR)]),
        "albba_gen_er": (C.c_int, [C.c_int32, C.c_double, C.c_uint64, C.c_uint64, C.POINTER(_CSR)]),
        "albba_gen_grid": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_uint64, C.c_uint64,
                                     C.POINTER(_CSR)]),
        "albba_pick_roots": (C.c_int, [C.POINTER(_CSR), C.c_int, C.c_uint64, C.c_int, I32P]),
        "albba_sellcs_build": (C.c_int, [C.POINTER(_CSR), C.c_int, C.c_int, C.c_int,
                                         C.POINTER(C.POINTER(_SELL))]),
        "albba_sellcs_build_ex": (C.c_int, [C.POINTER(_CSR), C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.POINTER(C.POINTER(_SELL))]),
        "albba_sellcs_free": (None, [C.POINTER(_SELL)]),
        "albba_device_count": (C.c_int, [C.POINTER(C.c_int)]),
        "albba_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
        "albba_ctx_destroy": (None, [C.c_void_p]),
        "albba_ctx_synchronize": (C.c_int, [C.c_void_p]),
        "albba_bench_copy": (C.c_int, [C.c_void_p, C.c_int64, C.c_int, C.POINTER(C.c_float)]),
        "albba_graph_upload": (C.c_int, [C.c_void_p, C.POINTER(_SELL), C.POINTER(C.c_void_p)]),
        "albba_graph_build": (C.c_int, [C.c_void_p, C.POINTER(_CSR), C.c_int, C.c_int, C.c_int,
                                        C.POINTER(C.c_void_p), C.POINTER(_BuildReport)]),
        "albba_graph_build_ex": (C.c_int, [C.c_void_p, C.POINTER(_CSR), C.c_int, C.c_int, C.c_int,
                                           C.c_int, C.POINTER(C.c_void_p), C.POINTER(_BuildReport)]),
        "albba_graph_layout": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.POINTER(_SELL))]),
        "albba_graph_generate": (C.c_int, [C.c_void_p, C.POINTER(_GenParams), C.c_int, C.c_int, C.c_int,
                                           C.POINTER(C.c_void_p), C.POINTER(_CSR),
                                           C.POINTER(_GenReport)]),
        "albba_graph_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
        "albba_graph_load": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(C.c_void_p), C.POINTER(_LoadReport)]),
        "albba_graph_free": (None, [C.c_void_p]),
        "albba_graph_device_bytes": (C.c_int64, [C.c_void_p]),
        "albba_bfs_opts_default": (None, [C.POINTER(_Opts)]),
        "albba_bfs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(_Opts),
                                C.POINTER(_Stats)]),
        "albba_bfs_multi": (C.c_int, [C.c_void_p, C.c_void_p, I32P, C.c_int32, C.POINTER(_Opts),
                                      C.c_int32, C.POINTER(_Stats)]),
        "albba_bfs_results": (C.c_int, [C.c_void_p, C.c_void_p, I32P, I32P]),
        "albba_selector_create": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
        "albba_batch_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int32,
                                         C.POINTER(C.c_void_p)]),
        "albba_batch_free": (None, [C.c_void_p]),
        "albba_batch_forget": (None, [C.c_void_p]),
        "albba_batch_run": (C.c_int, [C.c_void_p, I32P, C.POINTER(_Opts), I32P, C.c_int32,
                                      C.POINTER(_BatchResult), C.POINTER(C.c_float)]),
        "albba_batch_run_noalloc": (C.c_int, [C.c_void_p, I32P, C.POINTER(_Opts), I32P, C.c_int32,
                                              C.POINTER(_BatchResult), C.POINTER(C.c_float)]),
        "albba_batch_results": (C.c_int, [C.c_void_p, C.c_int32, I32P, I32P]),
        "albba_autotune": (C.c_int, [C.c_void_p, C.POINTER(_CSR), I32P, C.c_int32, C.POINTER(_Opts),
                                     C.POINTER(_TuneConfig), C.c_int32, C.POINTER(_TuneConfig),
                                     C.POINTER(_TuneConfig)]),
        "albba_selector_free": (None, [C.c_void_p]),
        "albba_selector_get_params": (C.c_int, [C.c_void_p, F32P]),
        "albba_selector_set_params": (C.c_int, [C.c_void_p, F32P]),
        "albba_selector_load": (C.c_int, [C.c_void_p, C.c_char_p]),
        "albba_selector_save": (C.c_int, [C.c_void_p, C.c_char_p]),
        "albba_selector_set_queue_ratio": (C.c_int, [C.c_void_p, C.c_float]),
        "albba_selector_queue_ratio": (C.c_float, [C.c_void_p]),
        "albba_selector_train": (C.c_int, [C.c_void_p, F32P, F32P, C.c_int32, C.c_int32, C.c_float,
                                           F32P]),
        "albba_selector_loss_grad": (C.c_int, [C.c_void_p, F32P, F32P, C.c_int32, F32P, F32P]),
        "albba_selector_predict": (C.c_int, [C.c_void_p, F32P, C.c_int32, F32P]),
        # reference-compatible entry points
        "read_Dmatrix": (None, [C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.POINTER(C.c_int)),
                                C.POINTER(F64P), C.POINTER(C.c_int), C.c_char_p]),
        "binary_read_csr": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.POINTER(C.c_int)),
                                      C.POINTER(F64P), C.c_char_p]),
        "binary_write_csr": (None, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int),
                                    C.POINTER(C.c_int), F64P, C.c_char_p]),
        "read_Dmatrix_convert": (None, [C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                        C.POINTER(C.POINTER(C.c_int)),
                                        C.POINTER(C.POINTER(C.c_int)), C.POINTER(F64P),
                                        C.POINTER(C.c_int), C.c_char_p]),
        "mmio_info": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                C.POINTER(C.c_int), C.c_char_p]),
        "mmio_data": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int), F64P, C.c_char_p]),
        "convert_csr_to_csc": (None, [C.POINTER(SparseMatrixCSR), C.POINTER(SparseMatrixCSC)]),
        "sellcs_preprocess": (None, [C.POINTER(SparseMatrixCSR), C.POINTER(SparseMatrixSELLCS),
                                     C.c_bool]),
        "free_sellcs_struct": (None, [C.POINTER(SparseMatrixSELLCS)]),
        "run_custom_bfs_test": (None, [C.POINTER(SparseMatrixCSR), F64P, C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       F64P, F64P, F64P, C.POINTER(C.c_int), C.c_bool, C.c_bool]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


_libc = C.CDLL(None)
_libc.free.argtypes = [C.c_void_p]
_libc.free.restype = None


def _check(rc: int) -> None:
    if rc != 0:
        raise AlbbaError(rc, lib().albba_last_error().decode(errors="replace"))


def _view(ptr, count, dtype):
    if count == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(int(count),)).view(dtype)


def rng_u64(seed: int, idx: int) -> int:
    return int(lib().albba_rng_u64(seed, idx))


# ---------------------------------------------------------------------------
# host graphs


class Graph:
    """A CSR working matrix (row v lists the in-neighbours of v)."""

    def __init__(self, raw: _CSR, owned: bool, keep=None):
        self._raw = raw
        self._owned = owned
        self._keep = keep

    @property
    def n(self) -> int:
        return int(self._raw.n)

    @property
    def nnz(self) -> int:
        return int(self._raw.nnz)

    @property
    def symmetric(self) -> bool:
        return bool(self._raw.symmetric)

    @property
    def row_ptr(self) -> np.ndarray:
        return _view(self._raw.row_ptr, self.n + 1, np.int64)

    @property
    def col(self) -> np.ndarray:
        return _view(self._raw.col, self.nnz, np.int32)

    def degrees(self) -> np.ndarray:
        return np.diff(self.row_ptr)

    def __del__(self):
        if getattr(self, "_owned", False) and _lib is not None:
            _lib.albba_csr_free(C.byref(self._raw))
            self._owned = False

    @classmethod
    def from_arrays(cls, row_ptr, col, symmetric: bool = False) -> "Graph":
        rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
        cc = np.ascontiguousarray(col, dtype=np.int32)
        raw = _CSR(int(rp.size - 1), int(cc.size), rp.ctypes.data_as(I64P), cc.ctypes.data_as(I32P),
                   int(symmetric))
        return cls(raw, owned=False, keep=(rp, cc))

    @classmethod
    def rmat(cls, scale, edgefactor=16, a=0.57, b=0.19, c=0.19, seed=1, perm_seed=2) -> "Graph":
        raw = _CSR()
        _check(lib().albba_gen_rmat(scale, edgefactor, a, b, c, seed, perm_seed, C.byref(raw)))
        return cls(raw, owned=True)

    @classmethod
    def er(cls, n, mean_degree, seed=1, perm_seed=2) -> "Graph":
        raw = _CSR()
        _check(lib().albba_gen_er(n, mean_degree, seed, perm_seed, C.byref(raw)))
        return cls(raw, owned=True)

    @classmethod
    def grid(cls, rows, cols, p_remove=0.1, seed=4, perm_seed=2) -> "Graph":
        raw = _CSR()
        _check(lib().albba_gen_grid(rows, cols, p_remove, seed, perm_seed, C.byref(raw)))
        return cls(raw, owned=True)

    @classmethod
    def read(cls, path: str) -> "Graph":
        raw = _CSR()
        _check(lib().albba_read_graph(path.encode(), C.byref(raw)))
        return cls(raw, owned=True)

    def write_cbd(self, path: str) -> None:
        _check(lib().albba_write_cbd(path.encode(), C.byref(self._raw)))

    def pick_roots(self, count: int, root_seed: int, largest_cc: bool = False) -> np.ndarray:
        out = np.zeros(max(count, 1), dtype=np.int32)
        k = lib().albba_pick_roots(C.byref(self._raw), count, root_seed, int(largest_cc),
                                   out.ctypes.data_as(I32P))
        if k < 0:
            _check(k)
        return out[:k].copy()

    def sellcs(self, C_: int = 64, sigma: int = 4096, pad: int = 32,
               sort_first: bool = False) -> "SellCS":
        """The device layout built on the host (albba_sellcs_build_ex);
        sort_first: the reference's -s 1 first-column row order."""
        p = C.POINTER(_SELL)()
        _check(lib().albba_sellcs_build_ex(C.byref(self._raw), C_, sigma, pad,
                                           SELL_SORT_FIRST_COL if sort_first else 0, C.byref(p)))
        return SellCS(p, self)


class SellCS:
    """Device-ready SELL-C-sigma layout built on the host."""

    def __init__(self, ptr, graph: Graph):
        self._p = ptr
        self._graph = graph

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None:
            _lib.albba_sellcs_free(self._p)
            self._p = None

    @property
    def raw(self) -> _SELL:
        return self._p.contents

    def arrays(self) -> dict:
        r = self.raw
        npos = r.nslices * r.C
        return {
            "row_order": _view(r.row_order, r.n, np.int32).copy(),
            "pos_of": _view(r.pos_of, r.n, np.int32).copy(),
            "slice_widths": _view(r.slice_widths, r.nslices, np.int32).copy(),
            "slice_pointers": _view(r.slice_pointers, r.nslices + 1, np.int64).copy(),
            "vop_pointers": _view(r.vop_pointers, r.nslices + 1, np.int64).copy(),
            "vop_lengths": _view(r.vop_lengths, r.nvops, np.uint8).copy(),
            "column_indices": _view(r.column_indices, r.padded_nnz, np.int32).copy(),
            "row_lengths": _view(r.row_lengths, npos, np.int32).copy(),
            "push_ptr": _view(r.push_ptr, npos + 1, np.int64).copy(),
            "push_col": _view(r.push_col, r.nnz, np.int32).copy(),
            "pull_ptr": _view(r.pull_ptr, npos + 1, np.int64).copy() if r.pull_ptr else None,
            "pull_col": _view(r.pull_col, r.nnz, np.int32).copy() if r.pull_col else None,
            "order_flags": int(r.order_flags),
        }


# ---------------------------------------------------------------------------
# device


def device_count() -> int:
    n = C.c_int(0)
    rc = lib().albba_device_count(C.byref(n))
    return int(n.value) if rc == 0 else 0


# contexts still open; closed (children first) when the interpreter exits,
# before it tears its modules down, so that nothing is left to the order in
# which module globals disappear
_open_contexts = weakref.WeakSet()


@atexit.register
def _close_open_contexts():
    for ctx in list(_open_contexts):
        ctx.close()


class Context:
    """One HIP device + stream.  Objects created on it (device graphs,
    selectors, batches) are closed before the context itself; a context still
    open when the interpreter exits is closed then."""

    def __init__(self, device: int = 0):
        self._p = C.c_void_p()
        self._children = []
        _check(lib().albba_ctx_create(device, C.byref(self._p)))
        _open_contexts.add(self)

    def _adopt(self, obj):
        self._children.append(weakref.ref(obj))

    def synchronize(self):
        _check(lib().albba_ctx_synchronize(self._p))

    def bench_copy(self, nbytes: int, reps: int = 10) -> float:
        """Mean ms of one device-to-device copy of nbytes (copy kernel)."""
        ms = C.c_float()
        _check(lib().albba_bench_copy(self._p, int(nbytes), int(reps), C.byref(ms)))
        return float(ms.value)

    def close(self):
        # newest first: a batch goes before the device graphs it runs on
        for ref in reversed(getattr(self, "_children", [])):
            obj = ref()
            if obj is not None:
                obj.close()
        self._children = []
        if getattr(self, "_p", None) and self._p.value and _lib is not None:
            _lib.albba_ctx_destroy(self._p)
            self._p = C.c_void_p()
        _open_contexts.discard(self)

    def __del__(self):
        self.close()


class DeviceGraph:
    def __init__(self, ctx: Context, sell: SellCS):
        """Upload a layout built on the host (Graph.sellcs)."""
        self.ctx = ctx
        self.n = sell.raw.n
        self.nnz = int(sell.raw.nnz)
        self.build_report = None
        self._p = C.c_void_p()
        _check(lib().albba_graph_upload(ctx._p, sell._p, C.byref(self._p)))
        ctx._adopt(self)

    @classmethod
    def build(cls, ctx: Context, graph: "Graph", C_: int = 64, sigma: int = 4096,
              pad: int = 1, sort_first: bool = False) -> "DeviceGraph":
        """Build the layout on the device from the CSR (albba_graph_build_ex):
        the same device arrays as
        DeviceGraph(ctx, graph.sellcs(C_, sigma, pad, sort_first))."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.n = graph.n
        self.nnz = int(graph.nnz)
        self._p = C.c_void_p()
        rep = _BuildReport()
        _check(lib().albba_graph_build_ex(ctx._p, C.byref(graph._raw), C_, sigma, pad,
                                          SELL_SORT_FIRST_COL if sort_first else 0, C.byref(self._p),
                                          C.byref(rep)))
        self.build_report = {"upload_ms": float(rep.upload_ms), "build_ms": float(rep.build_ms),
                             "rows_sorted": bool(rep.rows_sorted)}
        ctx._adopt(self)
        return self

    @classmethod
    def load(cls, ctx: Context, path: str, C_: int = 64, sigma: int = 4096,
             pad: int = 1) -> "DeviceGraph":
        """Load a .cbd file into device memory and prepare and build it there
        (albba_graph_load): the same device graph as
        DeviceGraph(ctx, Graph.read(path).sellcs(C_, sigma, pad))."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self._p = C.c_void_p()
        rep = _LoadReport()
        _check(lib().albba_graph_load(ctx._p, path.encode(), C_, sigma, pad, C.byref(self._p),
                                      C.byref(rep)))
        self.build_report = {"read_ms": float(rep.read_ms), "prep_ms": float(rep.prep_ms),
                             "build_ms": float(rep.build_ms), "total_ms": float(rep.total_ms)}
        ctx._adopt(self)
        n, nnz = C.c_int32(), C.c_int64()
        _check(lib().albba_graph_size(self._p, C.byref(n), C.byref(nnz)))
        self.n, self.nnz = int(n.value), int(nnz.value)
        return self

    @classmethod
    def generate(cls, ctx: Context, kind: str, C_: int = 64, sigma: int = 4096, pad: int = 1,
                 host_copy: bool = True, **kw):
        """Generate on the device (albba_graph_generate) the graph that
        Graph.rmat / Graph.er / Graph.grid would return, and build its layout
        there.  kind: "rmat" (scale, edgefactor=16, a=.57, b=.19, c=.19),
        "er" (n, mean_degree) or "grid" (rows, cols, p_remove=0.1); seed,
        perm_seed.  Returns (DeviceGraph, Graph or None)."""
        p = _GenParams()
        p.seed, p.perm_seed = kw.get("seed", 1), kw.get("perm_seed", 2)
        if kind == "rmat":
            p.kind, p.scale, p.edgefactor = GEN_RMAT, kw["scale"], kw.get("edgefactor", 16)
            p.a, p.b, p.c = kw.get("a", 0.57), kw.get("b", 0.19), kw.get("c", 0.19)
        elif kind == "er":
            p.kind, p.n, p.mean_degree = GEN_ER, kw["n"], kw["mean_degree"]
        elif kind == "grid":
            p.kind, p.rows, p.cols = GEN_GRID, kw["rows"], kw["cols"]
            p.p_remove = kw.get("p_remove", 0.1)
            p.seed = kw.get("seed", 4)
        else:
            raise ValueError(f"unknown kind {kind!r}")
        self = cls.__new__(cls)
        self.ctx = ctx
        self._p = C.c_void_p()
        raw = _CSR()
        rep = _GenReport()
        _check(lib().albba_graph_generate(ctx._p, C.byref(p), C_, sigma, pad, C.byref(self._p),
                                          C.byref(raw) if host_copy else None, C.byref(rep)))
        self.build_report = {"perm_ms": float(rep.perm_ms), "gen_ms": float(rep.gen_ms),
                             "build_ms": float(rep.build_ms), "total_ms": float(rep.total_ms)}
        ctx._adopt(self)
        n, nnz = C.c_int32(), C.c_int64()
        _check(lib().albba_graph_size(self._p, C.byref(n), C.byref(nnz)))
        self.n, self.nnz = int(n.value), int(nnz.value)
        return self, (Graph(raw, owned=True) if host_copy else None)

    def layout(self) -> "SellCS":
        """The device layout copied to the host (albba_graph_layout)."""
        p = C.POINTER(_SELL)()
        _check(lib().albba_graph_layout(self.ctx._p, self._p, C.byref(p)))
        return SellCS(p, None)

    @property
    def device_bytes(self) -> int:
        return int(lib().albba_graph_device_bytes(self._p))

    def close(self):
        if getattr(self, "_p", None) and self._p.value and _lib is not None:
            _lib.albba_graph_free(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        self.close()

    def bfs(self, root: int, policy: int = SWITCH_HEURISTIC, fixed_choice: int = CHOICE_PULL_SB,
            script=None, selector: "Selector" = None, time_levels: bool = False,
            timing_reps: int = 3, capacity: int = 0, results: bool = True,
            small_frontier_edges: int = None, small_frontier_blocks: int = 0) -> dict:
        """Run one BFS on the device; returns stats (+ level/parent arrays).

        small_frontier_edges: edge cap of the persistent small-frontier push
        levels (0 disables them; None keeps the library default);
        small_frontier_blocks: its workgroup count (0: default)."""
        o = _Opts()
        lib().albba_bfs_opts_default(C.byref(o))
        o.switch_policy = policy
        o.fixed_choice = fixed_choice
        o.time_levels = int(time_levels)
        o.timing_reps = timing_reps
        if small_frontier_edges is not None:
            o.small_frontier_edges = int(small_frontier_edges)
        o.small_frontier_blocks = int(small_frontier_blocks)
        scr = None
        if script is not None:
            scr = np.ascontiguousarray(script, dtype=np.int32)
            o.script = scr.ctypes.data_as(I32P)
            o.script_len = scr.size
        if selector is not None:
            o.selector = selector._p
        cap = capacity if capacity > 0 else 4096
        fr = np.zeros(cap, np.int32)
        ch = np.zeros(cap, np.int32)
        lm = np.zeros(cap, np.float32)
        ft = np.zeros(cap * NUM_FEATURES, np.float32)
        cm = np.zeros(cap * 4, np.float32)
        en = np.zeros(cap, np.int64)
        wd = np.zeros(cap, np.int64)
        by = np.zeros(cap, np.int64)
        km = np.zeros(cap, np.float32)
        cb = np.zeros(cap, np.int64)
        sw = np.full(cap, -1, np.int64)
        st = _Stats(0, cap, 0, 0, 0.0, fr.ctypes.data_as(I32P), ch.ctypes.data_as(I32P),
                    lm.ctypes.data_as(F32P), ft.ctypes.data_as(F32P), cm.ctypes.data_as(F32P),
                    en.ctypes.data_as(I64P), wd.ctypes.data_as(I64P), by.ctypes.data_as(I64P),
                    km.ctypes.data_as(F32P), cb.ctypes.data_as(I64P), sw.ctypes.data_as(I64P))
        _check(lib().albba_bfs(self.ctx._p, self._p, int(root), C.byref(o), C.byref(st)))
        L = min(st.num_levels, cap)
        out = {
            "num_levels": int(st.num_levels), "reached": int(st.reached),
            "edges_in_component": int(st.edges_in_component), "bfs_ms": float(st.bfs_ms),
            "frontier": fr[:L].copy(), "choice": ch[:L].copy(), "level_ms": lm[:L].copy(),
            "features": ft[:L * NUM_FEATURES].reshape(L, NUM_FEATURES).copy(),
            "choice_ms": cm[:L * 4].reshape(L, 4).copy(), "entries": en[:L].copy(),
            "words": wd[:L].copy(), "bypassed": by[:L].copy(), "kernel_ms": km[:L].copy(),
            "col_bypassed": cb[:L].copy(),
            # stored entries of the words a list pull scanned (time_levels)
            "swept": sw[:L].copy(),
        }
        if results:
            out["level"], out["parent"] = self.results()
        return out

    def bfs_many(self, roots, policy: int = SWITCH_HEURISTIC, selector: "Selector" = None,
                 scripts=None) -> dict:
        """BFS from each root in turn in one native call (albba_bfs_multi):
        each BFS is queued behind the previous one without a host
        synchronisation between them.  scripts[i]: root i's per-level
        choices (the scripted policy).  Returns per-root arrays; the graph's
        level/parent results are the last root's."""
        r = np.ascontiguousarray(roots, dtype=np.int32)
        k = r.size
        keep = []
        opts = (_Opts * (1 if scripts is None else k))()
        for i in range(len(opts)):
            o = opts[i]
            lib().albba_bfs_opts_default(C.byref(o))
            o.switch_policy = policy
            if selector is not None:
                o.selector = selector._p
            if scripts is not None:
                scr = np.ascontiguousarray(scripts[i], dtype=np.int32)
                keep.append(scr)
                o.script = scr.ctypes.data_as(I32P)
                o.script_len = scr.size
        st = (_Stats * max(k, 1))()
        _check(lib().albba_bfs_multi(self.ctx._p, self._p, r.ctypes.data_as(I32P), k, opts,
                                     len(opts), st))
        return {
            "bfs_ms": np.array([st[i].bfs_ms for i in range(k)]),
            "num_levels": np.array([st[i].num_levels for i in range(k)], np.int32),
            "reached": np.array([st[i].reached for i in range(k)], np.int64),
            "edges_in_component": np.array([st[i].edges_in_component for i in range(k)], np.int64),
        }

    def results(self):
        lv = np.empty(self.n, np.int32)
        pa = np.empty(self.n, np.int32)
        _check(lib().albba_bfs_results(self.ctx._p, self._p, lv.ctypes.data_as(I32P),
                                       pa.ctypes.data_as(I32P)))
        return lv, pa


MAP_LANE_ROW, MAP_WAVE_ROW = 0, 1


class Batch:
    """Batched BFS: one instance per entry of `graphs` (device graphs; one
    may appear several times), all run in one persistent launch."""

    def __init__(self, ctx: Context, graphs):
        self.ctx = ctx
        self.graphs = list(graphs)
        arr = (C.c_void_p * len(self.graphs))(*[dg._p.value for dg in self.graphs])
        self._p = C.c_void_p()
        _check(lib().albba_batch_create(ctx._p, arr, len(self.graphs), C.byref(self._p)))
        ctx._adopt(self)

    def close(self):
        if getattr(self, "_p", None) and self._p.value and _lib is not None:
            _lib.albba_batch_free(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        self.close()

    def run(self, roots, policy: int = SWITCH_HEURISTIC, selector: "Selector" = None,
            mapping=None, workgroups: int = 0, noalloc: bool = False) -> dict:
        """BFS from roots[i] on instance i; per-instance results as arrays.
        noalloc: albba_batch_run_noalloc (the run allocates no memory)."""
        n = len(self.graphs)
        r = np.ascontiguousarray(roots, dtype=np.int32)
        if r.size != n:
            raise ValueError("one root per instance")
        o = _Opts()
        lib().albba_bfs_opts_default(C.byref(o))
        o.switch_policy = policy
        if selector is not None:
            o.selector = selector._p
        mp = None
        if mapping is not None:
            mp = np.ascontiguousarray(np.broadcast_to(mapping, (n,)), dtype=np.int32)
        res = (_BatchResult * n)()
        ms = C.c_float()
        fn = lib().albba_batch_run_noalloc if noalloc else lib().albba_batch_run
        _check(fn(self._p, r.ctypes.data_as(I32P), C.byref(o),
                  mp.ctypes.data_as(I32P) if mp is not None else None, int(workgroups), res, C.byref(ms)))
        out = {"total_ms": float(ms.value)}
        for name, _ in _BatchResult._fields_:
            out[name] = np.array([getattr(x, name) for x in res])
        return out

    def forget(self):
        """Drop the previous run's levels and times: the next run splits the
        workgroups by the instances' shares of the stored entries."""
        lib().albba_batch_forget(self._p)

    def results(self, i: int):
        g = self.graphs[i]
        lv = np.empty(g.n, np.int32)
        pa = np.empty(g.n, np.int32)
        _check(lib().albba_batch_results(self._p, int(i), lv.ctypes.data_as(I32P),
                                         pa.ctypes.data_as(I32P)))
        return lv, pa


def autotune(ctx: Context, graph: "Graph", roots, policy: int = SWITCH_HEURISTIC,
             selector: "Selector" = None, candidates=None) -> dict:
    """Measure every (C, sigma) candidate (default: C in {32,64,128} x sigma in
    {C, 256, 4096, n}) on `roots`, then the batched pull mapping on the
    fastest layout.  Returns {"best": {...}, "results": [{...}, ...]}."""
    r = np.ascontiguousarray(roots, dtype=np.int32)
    o = _Opts()
    lib().albba_bfs_opts_default(C.byref(o))
    o.switch_policy = policy
    if selector is not None:
        o.selector = selector._p
    cand = None
    n = 12
    if candidates is not None:
        n = len(candidates)
        cand = (_TuneConfig * n)()
        for i, (c_, s_) in enumerate(candidates):
            cand[i].C, cand[i].sigma = int(c_), int(s_)
    res = (_TuneConfig * n)()
    best = _TuneConfig()
    _check(lib().albba_autotune(ctx._p, C.byref(graph._raw), r.ctypes.data_as(I32P), r.size,
                                C.byref(o), cand, n if cand is not None else 0, res, C.byref(best)))

    def as_dict(t):
        return {f: getattr(t, f) for f, _ in _TuneConfig._fields_}

    return {"best": as_dict(best), "results": [as_dict(t) for t in res]}


class Selector:
    """The learned direction selector (MLP 12 -> 32 -> 32 -> 4 on the GPU)."""

    def __init__(self, ctx: Context, seed: int = 1):
        self.ctx = ctx
        self._p = C.c_void_p()
        _check(lib().albba_selector_create(ctx._p, seed, C.byref(self._p)))
        ctx._adopt(self)

    def close(self):
        if getattr(self, "_p", None) and self._p.value and _lib is not None:
            _lib.albba_selector_free(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        self.close()

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(MLP_NPARAM, np.float32)
        _check(lib().albba_selector_get_params(self._p, p.ctypes.data_as(F32P)))
        return p

    @params.setter
    def params(self, v):
        v = np.ascontiguousarray(v, dtype=np.float32)
        assert v.size == MLP_NPARAM
        _check(lib().albba_selector_set_params(self._p, v.ctypes.data_as(F32P)))

    def load(self, path: str):
        _check(lib().albba_selector_load(self._p, path.encode()))

    @property
    def queue_ratio(self) -> float:
        """Push levels list the next queue while their frontier out-edges
        are at most queue_ratio * n (0: the heuristic's threshold)."""
        return float(lib().albba_selector_queue_ratio(self._p))

    @queue_ratio.setter
    def queue_ratio(self, r: float):
        _check(lib().albba_selector_set_queue_ratio(self._p, float(r)))

    def save(self, path: str):
        _check(lib().albba_selector_save(self._p, path.encode()))

    def train(self, x, t, steps: int = 2000, lr: float = 3e-3) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float(0)
        _check(lib().albba_selector_train(self._p, x.ctypes.data_as(F32P), t.ctypes.data_as(F32P),
                                          x.shape[0], steps, lr, C.byref(loss)))
        return float(loss.value)

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float(0)
        g = np.zeros(MLP_NPARAM, np.float32)
        _check(lib().albba_selector_loss_grad(self._p, x.ctypes.data_as(F32P),
                                              t.ctypes.data_as(F32P), x.shape[0], C.byref(loss),
                                              g.ctypes.data_as(F32P)))
        return float(loss.value), g

    def predict(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros((x.shape[0], 4), np.float32)
        _check(lib().albba_selector_predict(self._p, x.ctypes.data_as(F32P), x.shape[0],
                                            out.ctypes.data_as(F32P)))
        return out


# ---------------------------------------------------------------------------
# reference-compatible entry points


def read_Dmatrix(filename: str):
    """bfs/src/biio.h:127.  Returns (nrows, ncols, nnz, row_ptr, col_idx, val,
    isSymmetric) or None when nothing was loaded (unsupported extension or a
    failed read, where the reference prints and leaves its outputs unset)."""
    L = lib()
    r, c, z, s = C.c_int(-1), C.c_int(-1), C.c_int(-1), C.c_int(-1)
    rp, ci = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
    va = F64P()
    L.read_Dmatrix(C.byref(r), C.byref(c), C.byref(z), C.byref(rp), C.byref(ci), C.byref(va),
                   C.byref(s), filename.encode())
    if not rp or not ci or r.value < 0:
        return None
    out = (r.value, c.value, z.value, _view(rp, r.value + 1, np.int32).copy(),
           _view(ci, z.value, np.int32).copy(), _view(va, z.value, np.float64).copy(), s.value)
    for p in (rp, ci, va):
        _libc.free(C.cast(p, C.c_void_p))
    return out


def _csr_struct(n, row_ptr, col, name=b"graph"):
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    cc = np.ascontiguousarray(col, dtype=np.int32)
    vals = np.ones(max(cc.size, 1), dtype=np.float64)
    m = SparseMatrixCSR(name, vals.ctypes.data_as(F64P), cc.ctypes.data_as(C.POINTER(C.c_int)),
                        rp.ctypes.data_as(C.POINTER(C.c_int)), n, n, cc.size, 0)
    return m, (rp, cc, vals)


def convert_csr_to_csc(n, row_ptr, col):
    """bfs/src/sparse_matrix.c:230 on a square matrix: (col_ptr, row_idx)."""
    m, keep = _csr_struct(n, row_ptr, col)
    t = SparseMatrixCSC()
    lib().convert_csr_to_csc(C.byref(m), C.byref(t))
    cp = _view(t.column_pointers, n + 1, np.int32).copy()
    ri = _view(t.row_indices, m.nnz, np.int32).copy()
    for p in (t.column_pointers, t.row_indices, t.values):
        _libc.free(C.cast(p, C.c_void_p))
    return cp, ri


def sellcs_preprocess(n, row_ptr, col, C_=256, sigma=16384, sort_flag=False) -> dict:
    """bfs/src/bfs_SELLCSv2.c:34 -- the reference-compatible SELL layout."""
    m, keep = _csr_struct(n, row_ptr, col)
    s = SparseMatrixSELLCS()
    s.C = C_
    s.sigma = sigma
    lib().sellcs_preprocess(C.byref(m), C.byref(s), bool(sort_flag))
    nv = s.vop_pointers[s.nslices]
    out = {
        "nslices": s.nslices,
        "row_order": _view(s.row_order, n, np.int32).copy(),
        "slice_widths": _view(s.slice_widths, s.nslices, np.int32).copy(),
        "vop_pointers": _view(s.vop_pointers, s.nslices + 1, np.int32).copy(),
        "vop_lengths": _view(s.vop_lengths, nv, np.uint8).copy(),
        "slice_pointers": _view(s.slice_pointers, s.nslices + 1, np.int32).copy(),
        "column_indices": _view(s.column_indices, m.nnz, np.int32).copy(),
        "max_index": _view(s.max_index, nv, np.int32).copy(),
        "min_index": _view(s.min_index, nv, np.int32).copy(),
    }
    lib().free_sellcs_struct(C.byref(s))
    return out


def run_custom_bfs_test(n, row_ptr, col, sort_flag: bool = False) -> dict:
    """bfs/src/bfs.h:13 on the GPU (root 0): frontier sizes, iterations, times;
    sort_flag: the reference's -s 1 row order for the layout."""
    m, keep = _csr_struct(n, row_ptr, col)
    bfs_x = np.zeros(max(n, 1), np.float64)
    colidx_x = np.zeros(max(n, 1), np.int32)
    frontier = np.zeros(max(n, 1), np.int32)
    iter_times = np.zeros(max(n, 1), np.float64)
    nnzx, num_vex, num_iter = C.c_int(1), C.c_int(1), C.c_int(0)
    conv, bfs_t = C.c_double(0), C.c_double(0)
    lib().run_custom_bfs_test(C.byref(m), bfs_x.ctypes.data_as(F64P),
                              colidx_x.ctypes.data_as(C.POINTER(C.c_int)), C.byref(nnzx),
                              frontier.ctypes.data_as(C.POINTER(C.c_int)), C.byref(num_vex),
                              C.byref(conv), iter_times.ctypes.data_as(F64P), C.byref(bfs_t),
                              C.byref(num_iter), True, bool(sort_flag))
    k = num_iter.value
    return {"num_iter": k, "num_vex": num_vex.value, "frontier": frontier[:k].copy(),
            "iter_times": iter_times[:k].copy(), "bfs_time": bfs_t.value,
            "conv_time": conv.value, "nnzx": nnzx.value}


def gteps(nnz_stored: int, ms: float) -> float:
    """bfs/src/report_results.c:36: stored entries / (ms * 1e6)."""
    return nnz_stored / (ms * 1e6) if ms > 0 else float("nan")


def harmonic_mean(xs) -> float:
    xs = np.asarray(xs, dtype=np.float64)
    return float(len(xs) / np.sum(1.0 / xs)) if len(xs) else float("nan")
