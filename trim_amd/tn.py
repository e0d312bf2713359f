"""ctypes binding of libtn_hip.so (the C ABI declared in include/tn_hip.h).

Thin on purpose: every call goes straight to the HIP library, and a missing
library or device raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libtn_hip.so")

TN_GEMM_F64, TN_GEMM_F32, TN_GEMM_BF16X3, TN_GEMM_BF16X6, TN_GEMM_BF16_EXACT = 0, 1, 2, 3, 4
GEMM_MODES = {"f64": 0, "f32": 1, "bf16x3": 2, "bf16x6": 3, "bf16_exact": 4}

P = ctypes.c_void_p
I64 = ctypes.c_int64
PI64 = ctypes.POINTER(ctypes.c_int64)

# name -> argtypes, in the order of include/tn_hip.h
SIGNATURES = {
    "tn_last_error": [],
    "tn_init": [ctypes.c_int],
    "tn_device_info": [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_size_t),
                       ctypes.c_char_p, ctypes.c_int],
    "tn_set_budget": [ctypes.c_size_t],
    "tn_mem_stats": [ctypes.POINTER(ctypes.c_size_t)] * 4,
    "tn_malloc": [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t],
    "tn_free": [P],
    "tn_pool_trim": [],
    "tn_pool_set_stream": [P],
    "tn_memcpy_h2d": [P, P, ctypes.c_size_t, P],
    "tn_memcpy_d2h": [P, P, ctypes.c_size_t, P],
    "tn_memcpy_d2d": [P, P, ctypes.c_size_t, P],
    "tn_zero": [P, ctypes.c_size_t, P],
    "tn_stream_create": [ctypes.POINTER(ctypes.c_void_p)],
    "tn_stream_destroy": [P],
    "tn_stream_sync": [P],
    "tn_stream_query": [P, ctypes.POINTER(ctypes.c_int)],
    "tn_event_query": [P, ctypes.POINTER(ctypes.c_int)],
    "tn_host_alloc": [ctypes.POINTER(P), ctypes.c_size_t],
    "tn_host_free": [P],
    "tn_memcpy_d2h_async": [P, P, ctypes.c_size_t, P],
    "tn_device_sync": [],
    "tn_event_create": [ctypes.POINTER(ctypes.c_void_p)],
    "tn_event_destroy": [P],
    "tn_event_record": [P, P],
    "tn_event_elapsed_ms": [P, P, ctypes.POINTER(ctypes.c_float)],
    "tn_graph_begin": [P],
    "tn_graph_end": [P, ctypes.POINTER(ctypes.c_void_p)],
    "tn_graph_launch": [P, P],
    "tn_graph_destroy": [P],
    "tn_permute_dense_f64": [P, P, ctypes.c_int, P, P, P],
    "tn_reshape_elem_f64": [P, ctypes.c_int, ctypes.c_size_t, P, P, P],
    "tn_blockperm_plan_create": [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, I64, P, P, P, P, I64],
    "tn_blockperm_exec": [P, P, P, P],
    "tn_blockperm_plan_destroy": [P],
    "tn_rectscale_plan_create": [ctypes.POINTER(ctypes.c_void_p), I64, P],
    "tn_rectscale_exec": [P, P, ctypes.c_double, P],
    "tn_rectscale_plan_destroy": [P],
    "tn_transpose_f64": [P, I64, I64, P, P],
    "tn_gemm_plan_create": [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, I64, P],
    "tn_gemm_plan_create_ex": [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, I64, P],
    "tn_gemm_plan_set_tile": [P, ctypes.c_int],
    "tn_gemm_plan_autotune": [P, P, P, P, ctypes.c_int, P, ctypes.POINTER(ctypes.c_float)],
    "tn_gemm_plan_info": [P, PI64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)],
    "tn_gemm_plan_set_mask": [P, P, P],
    "tn_gemm_exec": [P, P, P, P, P],
    "tn_gemm_plan_destroy": [P],
    "tn_matmul_f64": [P, P, I64, I64, I64, P, ctypes.c_int, P],
    "tn_permute_dense_c128": [P, P, ctypes.c_int, P, P, P],
    "tn_matmul_c128": [P, P, I64, I64, I64, P, P],
    "tn_norm2_f64": [P, I64, P, P],
    "tn_sum_f64": [P, I64, I64, P, P],
    "tn_fill_uniform_f64": [P, I64, ctypes.c_uint64, P],
    "tn_axpy_f64": [ctypes.c_double, P, P, I64, P],
    "tn_copy_strided_f64": [P, I64, P, I64, I64, P],
    "tn_scale_f64": [P, I64, ctypes.c_double, P],
    "tn_scale_inv_dev_f64": [P, I64, P, P],
    "tn_diag_row_mul_f64": [P, P, I64, I64, P],
    "tn_diag_col_mul_f64": [P, P, I64, I64, P],
    "tn_svd_f64": [P, I64, I64, P, P, P, P],
    "tn_svd_f64_noalloc": [P, I64, I64, P, P, P, P],
    "tn_svd_f64_batched": [I64, P, P, P, P, P, P, P, P, P, P, P],
    "tn_svd_last_sweeps": [ctypes.POINTER(ctypes.c_int)],
    "tn_svd_set_block_min": [ctypes.c_int64],
    "tn_svd_set_max_sweeps": [ctypes.c_int],
    "tn_svd_f64_warm": [P, I64, I64, P, P, P, P, P],
    "tn_svd_plan_create": [ctypes.POINTER(ctypes.c_void_p), I64, P, P, P, P, P, P, P],
    "tn_svd_plan_exec": [P, P, P, P, P, P, P],
    "tn_svd_plan_check": [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), P],
    "tn_svd_plan_set_mask": [P, P, P],
    "tn_svd_plan_destroy": [P],
    "tn_vecops_plan_create": [ctypes.POINTER(ctypes.c_void_p), I64, P],
    "tn_vecops_plan_exec": [P, P, P],
    "tn_vecops_plan_set_mask": [P, P, P],
    "tn_vecops_plan_destroy": [P],
    "tn_eigh_f64": [P, I64, P, P, P],
    "tn_eigh_f64_batched": [I64, PI64, P, PI64, P, PI64, P, PI64, P],
    "tn_svd_c128": [P, I64, I64, P, P, P, P],
    "tn_eigh_c128": [P, I64, P, P, P],
    "tn_expand_c128": [P, I64, P, P],
    "tn_lanczos_f64": [P, I64, P, PI64, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                       ctypes.POINTER(ctypes.c_int), P],
    "tn_lanczos_c128": [P, I64, P, PI64, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                        ctypes.POINTER(ctypes.c_int), P],
    "tn_truncate_f64": [P, I64, I64, P, P, P, P],
}

_lib = None


class TNError(RuntimeError):
    pass


def load(path: Optional[str] = None):
    """Load libtn_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise TNError("libtn_hip.so is not built (%s); run python -m trim_amd._build" % path)
    L = ctypes.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_char_p if name == "tn_last_error" else ctypes.c_int
    _lib = L
    return L


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = load().tn_last_error().decode(errors="replace")
        raise TNError("%s failed (rc=%d): %s" % (what, rc, msg))


def call(name: str, *args):
    rc = getattr(load(), name)(*args)
    check(rc, name)
    return rc


def init(device: int = 0):
    call("tn_init", device)


def device_info():
    cus = ctypes.c_int()
    hbm = ctypes.c_size_t()
    arch = ctypes.create_string_buffer(64)
    call("tn_device_info", ctypes.byref(cus), ctypes.byref(hbm), arch, 64)
    return {"cu_count": cus.value, "hbm_bytes": hbm.value, "arch": arch.value.decode()}


def _p(a) -> ctypes.c_void_p:
    if a is None:
        return ctypes.c_void_p(0)
    if isinstance(a, DeviceArray):
        return ctypes.c_void_p(a.ptr)
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(ctypes.c_void_p)
    if isinstance(a, int):
        return ctypes.c_void_p(a)
    return a


class DeviceArray:
    """A device buffer from the library's pool, with a numpy dtype and shape."""

    def __init__(self, shape, dtype=np.float64, zero: bool = False):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        ptr = ctypes.c_void_p()
        call("tn_malloc", ctypes.byref(ptr), max(self.nbytes, 8))
        self.ptr = ptr.value
        if zero:
            call("tn_zero", ctypes.c_void_p(self.ptr), max(self.nbytes, 8), None)

    @classmethod
    def from_host(cls, a: np.ndarray, stream=None) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if a.nbytes:
            call("tn_memcpy_h2d", ctypes.c_void_p(d.ptr), a.ctypes.data_as(ctypes.c_void_p),
                 a.nbytes, stream)
        return d

    def to_host(self, stream=None) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            call("tn_memcpy_d2h", out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr),
                 self.nbytes, stream)
        return out

    def offset(self, elems: int) -> int:
        return self.ptr + elems * self.dtype.itemsize

    def free(self):
        if self.ptr:
            load().tn_free(ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def sync(stream=None):
    call("tn_stream_sync", stream)


# --- thin op wrappers ---------------------------------------------------------
def permute_dense(src: DeviceArray, dst: DeviceArray, dims: Sequence[int], perm: Sequence[int],
                  stream=None):
    d = np.asarray(dims, dtype=np.int64)
    p = np.asarray(perm, dtype=np.int32)
    call("tn_permute_dense_f64", _p(src), _p(dst), len(d), _p(d), _p(p), stream)


def reshape_elem(src: DeviceArray, dst: DeviceArray, offset: Sequence[int], stream=None):
    off = np.asarray(offset, dtype=np.uint64)
    r = len(off) // 2
    call("tn_reshape_elem_f64", _p(src), r, src.size, _p(off), _p(dst), stream)


class BlockPermPlan:
    def __init__(self, rank, meta, sub_dims, bot_acc, sign, dst_elems):
        self.meta = np.ascontiguousarray(meta, dtype=np.int64)
        self.sub = np.ascontiguousarray(sub_dims, dtype=np.int64)
        self.acc = np.ascontiguousarray(bot_acc, dtype=np.int64)
        self.sign = np.ascontiguousarray(sign, dtype=np.float64)
        h = ctypes.c_void_p()
        call("tn_blockperm_plan_create", ctypes.byref(h), rank, self.meta.shape[0], _p(self.meta),
             _p(self.sub), _p(self.acc), _p(self.sign), int(dst_elems))
        self.h = h.value

    def __call__(self, src, dst, stream=None):
        call("tn_blockperm_exec", ctypes.c_void_p(self.h), _p(src), _p(dst), stream)

    def __del__(self):
        try:
            if self.h:
                load().tn_device_sync()
                load().tn_blockperm_plan_destroy(ctypes.c_void_p(self.h))
                self.h = 0
        except Exception:
            pass


GEMM_DESC = np.dtype([("M", "<i8"), ("N", "<i8"), ("K", "<i8"), ("a_off", "<i8"),
                      ("b_off", "<i8"), ("c_off", "<i8"), ("lda", "<i8"), ("ldb", "<i8"),
                      ("ldc", "<i8")])


INDEX_MAP = np.dtype([("n", "<i4"), ("pad", "<i4"), ("ext", "<i8", (4,)), ("stride", "<i8", (4,))])
GEMM_DESC_EX = np.dtype([("M", "<i8"), ("N", "<i8"), ("K", "<i8"), ("a_off", "<i8"), ("b_off", "<i8"),
                         ("c_off", "<i8"), ("lda", "<i8"), ("ldb", "<i8"), ("ldc", "<i8"),
                         ("a_row", INDEX_MAP), ("a_col", INDEX_MAP), ("b_row", INDEX_MAP),
                         ("b_col", INDEX_MAP), ("c_row", INDEX_MAP), ("c_col", INDEX_MAP),
                         ("d_off", "<i8"), ("rs_off", "<i8"), ("rs_div", "<i8")])


def index_map(levels):
    """[(extent, stride), ...] outermost first -> an INDEX_MAP record (empty = plain)."""
    m = np.zeros((), dtype=INDEX_MAP)
    m["n"] = len(levels)
    for i, (e, s) in enumerate(levels):
        m["ext"][i] = e
        m["stride"][i] = s
    return m


class GemmPlan:
    def __init__(self, mode, descs):
        descs = np.asarray(descs)
        if descs.dtype == GEMM_DESC_EX:
            self.descs = np.ascontiguousarray(descs)
            fn = "tn_gemm_plan_create_ex"
        else:
            self.descs = np.ascontiguousarray(descs, dtype=GEMM_DESC)
            fn = "tn_gemm_plan_create"
        h = ctypes.c_void_p()
        call(fn, ctypes.byref(h), int(mode), len(self.descs), _p(self.descs))
        self.h = h.value

    def set_tile(self, tile_id: int):
        call("tn_gemm_plan_set_tile", ctypes.c_void_p(self.h), int(tile_id))

    def autotune(self, A, B, C, reps: int = 5, stream=None) -> float:
        ms = ctypes.c_float()
        call("tn_gemm_plan_autotune", ctypes.c_void_p(self.h), _p(A), _p(B), _p(C), int(reps), stream,
             ctypes.byref(ms))
        return ms.value

    def info(self):
        t = ctypes.c_int64()
        f = ctypes.c_double()
        c = ctypes.c_int()
        call("tn_gemm_plan_info", ctypes.c_void_p(self.h), ctypes.byref(t), ctypes.byref(f),
             ctypes.byref(c))
        return {"tiles": t.value, "flops": f.value, "tile_id": c.value}

    def __call__(self, A, B, C, stream=None):
        call("tn_gemm_exec", ctypes.c_void_p(self.h), _p(A), _p(B), _p(C), stream)

    def __del__(self):
        try:
            if self.h:
                load().tn_device_sync()
                load().tn_gemm_plan_destroy(ctypes.c_void_p(self.h))
                self.h = 0
        except Exception:
            pass


def svd(A: DeviceArray, M: int, N: int):
    k = min(M, N)
    U = DeviceArray((M, k))
    S = DeviceArray((k,))
    VT = DeviceArray((k, N))
    call("tn_svd_f64", _p(A), M, N, _p(U), _p(S), _p(VT), None)
    return U, S, VT


def eigh(A: DeviceArray, N: int):
    w = DeviceArray((N,))
    v = DeviceArray((N, N))
    call("tn_eigh_f64", _p(A), N, _p(w), _p(v), None)
    return w, v


class Graph:
    """Captures the library calls made by fn() on `stream` into a HIP graph."""

    def __init__(self, fn, stream=None):
        self.stream = stream
        call("tn_graph_begin", stream)
        try:
            fn()
        finally:
            h = ctypes.c_void_p()
            rc = load().tn_graph_end(stream, ctypes.byref(h))
        check(rc, "tn_graph_end")
        self.h = h.value

    def launch(self):
        call("tn_graph_launch", ctypes.c_void_p(self.h), self.stream)

    def __del__(self):
        try:
            if self.h:
                load().tn_graph_destroy(ctypes.c_void_p(self.h))
                self.h = 0
        except Exception:
            pass


def stream_create():
    h = ctypes.c_void_p()
    call("tn_stream_create", ctypes.byref(h))
    return h.value


def pool_set_stream(stream):
    """The stream this thread's allocations and frees are ordered on."""
    call("tn_pool_set_stream", stream)


class Event:
    def __init__(self):
        h = ctypes.c_void_p()
        call("tn_event_create", ctypes.byref(h))
        self.h = h.value

    def record(self, stream=None):
        call("tn_event_record", ctypes.c_void_p(self.h), stream)

    def query(self) -> bool:
        """True once the work recorded before the event is done (never blocks)."""
        done = ctypes.c_int()
        call("tn_event_query", ctypes.c_void_p(self.h), ctypes.byref(done))
        return bool(done.value)

    def destroy(self):
        if self.h:
            call("tn_event_destroy", ctypes.c_void_p(self.h))
            self.h = None

    def elapsed_ms(self, later: "Event") -> float:
        ms = ctypes.c_float()
        call("tn_event_elapsed_ms", ctypes.c_void_p(self.h), ctypes.c_void_p(later.h),
             ctypes.byref(ms))
        return ms.value
