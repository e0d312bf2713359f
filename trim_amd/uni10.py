"""Python binding of the Uni10 C++ API (libuni10_amd.so through include/uni10_capi.h).

Plays the role of the reference's SWIG module (uni10/Swig/pyUni10.i): the
classes hold handles of C++ objects whose elements live in HBM; every
operation runs the C++ API and its HIP kernels.  No computation happens in
Python and nothing falls back to a CPU path.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Sequence

import numpy as np

from . import tn as _tn

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libuni10_amd.so")
NETDIR = os.path.join(_PKG, "nets")     # the iTEBD drivers' network files

BD_IN, BD_OUT = 1, -1
PRT_EVEN, PRT_ODD = 0, 1
PRTF_EVEN, PRTF_ODD = 0, 1
GEMM_F64, GEMM_F32, GEMM_BF16X3, GEMM_BF16X6, GEMM_BF16_EXACT = 0, 1, 2, 3, 4

P = ctypes.c_void_p
PP = ctypes.POINTER(ctypes.c_void_p)
I = ctypes.c_int
I64 = ctypes.c_int64
D = ctypes.c_double
PI = ctypes.POINTER(ctypes.c_int)
PI64 = ctypes.POINTER(ctypes.c_int64)
PD = ctypes.POINTER(ctypes.c_double)
S = ctypes.c_char_p

SIGNATURES = {
    "u10_last_error": [],
    "u10_set_gemm_mode": [I],
    "u10_set_stream": [P],
    "u10_set_autotune": [I],
    "u10_tensor_create": [I, P, P, P, P, S, PP],
    "u10_tensor_create_c": [I, P, P, P, P, S, PP],
    "u10_tensor_type": [P, ctypes.POINTER(ctypes.c_int)],
    "u10_svd_c128": [P, I64, I64, P, P, P],
    "u10_eigh_c128": [P, I64, P, P],
    "u10_bond_combine": [I, P, P, P, P, I, I, PI, PI, PI, P, P, PI, P, P, I],
    "u10_tensor_scalar": [D, PP],
    "u10_tensor_load": [S, PP],
    "u10_tensor_save": [P, S],
    "u10_tensor_copy": [P, PP],
    "u10_tensor_free": [P],
    "u10_tensor_info": [P, PI, PI, PI64, PI],
    "u10_tensor_labels": [P, P],
    "u10_tensor_set_label": [P, I, P],
    "u10_tensor_bond": [P, I, PI, PI, PI],
    "u10_tensor_bond_runs": [P, I, P, P],
    "u10_tensor_block": [P, I, P, PI64, PI64, PI64],
    "u10_tensor_set_elem": [P, P],
    "u10_tensor_get_elem": [P, P],
    "u10_tensor_set_raw": [P, P],
    "u10_tensor_raw_shape": [P, PI64, PI64],
    "u10_tensor_get_raw": [P, P],
    "u10_tensor_at": [P, P, PD],
    "u10_tensor_put_block": [P, P, I64, I64, P],
    "u10_tensor_get_block": [P, P, P],
    "u10_tensor_permute": [P, I, P, I],
    "u10_tensor_transpose": [P],
    "u10_tensor_scale": [P, D],
    "u10_tensor_randomize": [P, ctypes.c_uint64],
    "u10_tensor_norm": [P, PD],
    "u10_tensor_device_ptr": [P, PP],
    "u10_contract": [P, P, I, PP],
    "u10_otimes": [P, P, PP],
    "u10_tensor_mul": [P, P, PP],
    "u10_tensor_ex_swap": [P, P, P, I, PI],
    "u10_tensor_add_gate": [P, I, P],
    "u10_network_print": [P, ctypes.c_char_p, I],
    "u10_network_set_graph": [P, I],
    "u10_network_create": [S, PP],
    "u10_network_put": [P, S, P, I],
    "u10_network_launch": [P, S, PP],
    "u10_network_tree": [P, ctypes.c_char_p, I],
    "u10_network_free": [P],
    "u10_rdotr": [I64, I64, I, P, I64, I64, I, P, PI64, PI64, PI, P, I64],
    "u10_svd": [P, I64, I64, P, P, P],
    "u10_eigh": [P, I64, P, P],
    "u10_lanczos_eigh": [P, I64, P, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                         ctypes.POINTER(ctypes.c_uint64)],
    "u10_exph": [D, P, I64, P],
    "u10_itebd_create": [I, I, D, D, ctypes.c_uint64, S, PP],
    "u10_itebd_step": [P, I],
    "u10_itebd_set_graph": [P, I],
    "u10_itebd_energy": [P, PD],
    "u10_itebd_lambda": [P, I, P, PI],
    "u10_itebd_trunc_err": [P, PD],
    "u10_itebd_free": [P],
    "u10_batch_tebd_create": [I, P, P, P, D, I, S, PP],
    "u10_batch_tebd_create_ex": [I, P, P, P, P, P, D, I, S, PP],
    "u10_batch_tebd_sectors": [P, I, I, P, P, PI],
    "u10_u1_bond_sectors": [I, I, I, P, P, PI],
    "u10_batch_tebd_step": [P, I],
    "u10_batch_tebd_launch": [P, I],
    "u10_batch_tebd_wait": [P],
    "u10_batch_tebd_ready": [P, PI],
    "u10_batch_tebd_set_graph": [P, I],
    "u10_batch_tebd_set_stream": [P, P],
    "u10_batch_tebd_set_active": [P, I, I],
    "u10_batch_tebd_load": [P, I, D, ctypes.c_uint64],
    "u10_batch_tebd_energy": [P, I, PD],
    "u10_batch_tebd_lambda": [P, I, I, P, PI],
    "u10_batch_tebd_trunc_err": [P, I, PD],
    "u10_batch_tebd_info": [P, PD, PI],
    "u10_batch_tebd_free": [P],
}

_lib = None


class Uni10Error(RuntimeError):
    pass


def load():
    global _lib
    if _lib is None:
        _tn.load()   # dependency first (same directory, rpath $ORIGIN)
        if not os.path.exists(LIB_PATH):
            raise Uni10Error("libuni10_amd.so is not built; run python -m trim_amd._build")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_char_p if name == "u10_last_error" else ctypes.c_int
        _lib = L
    return _lib


def _call(name, *args):
    rc = getattr(load(), name)(*args)
    if rc != 0:
        raise Uni10Error("%s: %s" % (name, load().u10_last_error().decode(errors="replace")))


def _arr(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data_as(ctypes.c_void_p)


def set_gemm_mode(mode: int):
    _call("u10_set_gemm_mode", int(mode))


def set_stream(stream):
    """The stream later calls of this thread enqueue on (None: the library's)."""
    _call("u10_set_stream", stream)


def set_autotune(on: bool):
    _call("u10_set_autotune", 1 if on else 0)


class Qnum:
    __slots__ = ("U1", "prt", "prtF")

    def __init__(self, U1=0, prt=PRT_EVEN, prtF=PRTF_EVEN):
        self.U1, self.prt, self.prtF = int(U1), int(prt), int(prtF)

    def triple(self):
        return (self.U1, self.prt, self.prtF)

    def __eq__(self, o):
        return self.triple() == o.triple()

    def __hash__(self):
        return hash(self.triple())

    def __repr__(self):
        return "Qnum(%d,%d,%d)" % self.triple()


class Bond:
    """Bond descriptor: a type and one Qnum per state (combine/degeneracy are
    C++ operations; this is what crosses the binding)."""

    def __init__(self, tp, qnums_or_dim):
        self.type = int(tp)
        if isinstance(qnums_or_dim, (int, np.integer)):
            self.qnums = [Qnum(0)] * int(qnums_or_dim)
        else:
            self.qnums = [q if isinstance(q, Qnum) else Qnum(*q) for q in qnums_or_dim]

    @property
    def dim(self):
        return len(self.qnums)

    def runs(self):
        out = []
        for q in self.qnums:
            if out and out[-1][0] == q:
                out[-1][1] += 1
            else:
                out.append([q, 1])
        return out


def combine_bonds(bonds: Sequence[Bond], tp: Optional[int] = None):
    """Bond::combine of the C++ API: bonds[0].combine(bonds[1])... (tp None),
    or the free combine(tp, bonds).  Returns (type, dim, runs, degeneracy)
    with runs and degeneracy as [(Qnum, count)]."""
    types = np.array([b.type for b in bonds], np.int32)
    runs = [b.runs() for b in bonds]
    nruns = np.array([len(r) for r in runs], np.int32)
    q = np.array([x[0].triple() for r in runs for x in r], np.int32).reshape(-1)
    d = np.array([x[1] for r in runs for x in r], np.int32)
    cap = 4096
    oq, od = np.zeros(3 * cap, np.int32), np.zeros(cap, np.int32)
    gq, gd = np.zeros(3 * cap, np.int32), np.zeros(cap, np.int32)
    ot, odim, onr, ong = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _call("u10_bond_combine", len(bonds), types.ctypes.data_as(P), nruns.ctypes.data_as(P), q.ctypes.data_as(P),
          d.ctypes.data_as(P), 0 if tp is None else 1, int(tp or 1), ctypes.byref(ot), ctypes.byref(odim),
          ctypes.byref(onr), oq.ctypes.data_as(P), od.ctypes.data_as(P), ctypes.byref(ong), gq.ctypes.data_as(P),
          gd.ctypes.data_as(P), cap)
    rl = [(Qnum(*oq[3 * i:3 * i + 3]), int(od[i])) for i in range(onr.value)]
    dg = [(Qnum(*gq[3 * i:3 * i + 3]), int(gd[i])) for i in range(ong.value)]
    return ot.value, odim.value, rl, dg


def _cplx(a):
    """complex array -> interleaved (re, im) float64, the device layout"""
    return np.ascontiguousarray(a, dtype=np.complex128).view(np.float64)


class UniTensor:
    """UniTensor (UniTensor.h).  dtype=complex makes a complex tensor
    (UniTensor(CTYPE, bonds)); its elements, raw elements and blocks come
    back as complex arrays."""

    def __init__(self, bonds: Sequence[Bond] = None, name: str = "", _handle=None, dtype=float):
        if _handle is not None:
            self.h = _handle
            return
        bonds = list(bonds or [])
        types = np.array([b.type for b in bonds], np.int32)
        runs = [b.runs() for b in bonds]
        nruns = np.array([len(r) for r in runs], np.int32)
        q = np.array([x[0].triple() for r in runs for x in r], np.int32).reshape(-1)
        d = np.array([x[1] for r in runs for x in r], np.int32)
        h = ctypes.c_void_p()
        fn = "u10_tensor_create_c" if np.dtype(dtype).kind == "c" else "u10_tensor_create"
        _call(fn, len(bonds), types.ctypes.data_as(P), nruns.ctypes.data_as(P),
              (q if q.size else np.zeros(1, np.int32)).ctypes.data_as(P),
              (d if d.size else np.zeros(1, np.int32)).ctypes.data_as(P), name.encode(), ctypes.byref(h))
        self.h = h.value

    def typeID(self) -> int:
        """1 real, 2 complex"""
        t = ctypes.c_int()
        _call("u10_tensor_type", self.h, ctypes.byref(t))
        return t.value

    @property
    def is_complex(self) -> bool:
        return self.typeID() == 2

    @classmethod
    def scalar(cls, v: float):
        h = ctypes.c_void_p()
        _call("u10_tensor_scalar", float(v), ctypes.byref(h))
        return cls(_handle=h.value)

    @classmethod
    def load(cls, fname: str):
        h = ctypes.c_void_p()
        _call("u10_tensor_load", fname.encode(), ctypes.byref(h))
        return cls(_handle=h.value)

    def save(self, fname: str):
        _call("u10_tensor_save", self.h, fname.encode())

    def copy(self):
        h = ctypes.c_void_p()
        _call("u10_tensor_copy", self.h, ctypes.byref(h))
        return UniTensor(_handle=h.value)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                load().u10_tensor_free(self.h)
                self.h = None
        except Exception:
            pass

    def _info(self):
        nb, ib, nblk = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ne = ctypes.c_int64()
        _call("u10_tensor_info", self.h, ctypes.byref(nb), ctypes.byref(ib), ctypes.byref(ne),
              ctypes.byref(nblk))
        return nb.value, ib.value, ne.value, nblk.value

    def bondNum(self):
        return self._info()[0]

    def inBondNum(self):
        return self._info()[1]

    def elemNum(self):
        return self._info()[2]

    def blockNum(self):
        return self._info()[3]

    def label(self) -> List[int]:
        n = self.bondNum()
        out = np.zeros(max(n, 1), np.int32)
        _call("u10_tensor_labels", self.h, out.ctypes.data_as(P))
        return [int(v) for v in out[:n]]

    def setLabel(self, labels):
        a, p = _arr(labels, np.int32)
        _call("u10_tensor_set_label", self.h, len(a), p)
        return self

    def bond(self, b=None):
        if b is None:
            return [self.bond(i) for i in range(self.bondNum())]
        tp, dim, nr = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _call("u10_tensor_bond", self.h, b, ctypes.byref(tp), ctypes.byref(dim), ctypes.byref(nr))
        q = np.zeros(3 * nr.value, np.int32)
        d = np.zeros(nr.value, np.int32)
        _call("u10_tensor_bond_runs", self.h, b, q.ctypes.data_as(P), d.ctypes.data_as(P))
        qs = []
        for i in range(nr.value):
            qs += [Qnum(*q[3 * i:3 * i + 3])] * int(d[i])
        return Bond(tp.value, qs)

    def blocks(self):
        """[(Qnum, rows, cols, offset)] in block order."""
        out = []
        for i in range(self.blockNum()):
            q = np.zeros(3, np.int32)
            r, c, o = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
            _call("u10_tensor_block", self.h, i, q.ctypes.data_as(P), ctypes.byref(r), ctypes.byref(c),
                  ctypes.byref(o))
            out.append((Qnum(*q), r.value, c.value, o.value))
        return out

    def setElem(self, e):
        if self.is_complex:
            a = _cplx(np.ravel(e))
            if a.size < 2 * self.elemNum():
                raise ValueError("not enough elements")
            _call("u10_tensor_set_elem", self.h, a.ctypes.data_as(P))
            return
        a, p = _arr(e, np.float64)
        if a.size < self.elemNum():
            raise ValueError("not enough elements")
        _call("u10_tensor_set_elem", self.h, p)

    def getElem(self) -> np.ndarray:
        cx = self.is_complex
        out = np.empty(self.elemNum() * (2 if cx else 1))
        _call("u10_tensor_get_elem", self.h, out.ctypes.data_as(P))
        return out.view(np.complex128) if cx else out

    def setRawElem(self, raw):
        if self.is_complex:
            a = _cplx(np.ravel(raw))
            _call("u10_tensor_set_raw", self.h, a.ctypes.data_as(P))
            return
        a, p = _arr(raw, np.float64)
        _call("u10_tensor_set_raw", self.h, p)

    def getRawElem(self) -> np.ndarray:
        r, c = ctypes.c_int64(), ctypes.c_int64()
        _call("u10_tensor_raw_shape", self.h, ctypes.byref(r), ctypes.byref(c))
        cx = self.is_complex
        out = np.empty((r.value, c.value * (2 if cx else 1)))
        _call("u10_tensor_get_raw", self.h, out.ctypes.data_as(P))
        return out.view(np.complex128) if cx else out

    def at(self, idxs):
        a, p = _arr(idxs, np.int64)
        v = ctypes.c_double()
        _call("u10_tensor_at", self.h, p, ctypes.byref(v))
        return v.value

    def putBlock(self, mat, q: Optional[Qnum] = None):
        q = q or Qnum(0)
        qa, qp = _arr(q.triple(), np.int32)
        if self.is_complex:
            m = np.ascontiguousarray(mat, dtype=np.complex128)
            _call("u10_tensor_put_block", self.h, qp, m.shape[0], m.shape[1], _cplx(m).ctypes.data_as(P))
            return
        m, p = _arr(mat, np.float64)
        _call("u10_tensor_put_block", self.h, qp, m.shape[0], m.shape[1], p)

    def getBlock(self, q: Optional[Qnum] = None) -> np.ndarray:
        q = q or Qnum(0)
        cx = self.is_complex
        for bq, r, c, o in self.blocks():     # (a complex block holds 2 doubles per element)
            if bq == q:
                out = np.empty((r, c))
                qa, qp = _arr(q.triple(), np.int32)
                _call("u10_tensor_get_block", self.h, qp, out.ctypes.data_as(P))
                return out.view(np.complex128) if cx else out
        raise KeyError("There is no block with the given quantum number %r" % (q,))

    def permute(self, labels, in_bond_num=None):
        if in_bond_num is None:          # permute(rowBondNum)
            in_bond_num, labels = labels, self.label()
        a, p = _arr(labels, np.int32)
        _call("u10_tensor_permute", self.h, len(a), p, int(in_bond_num))
        return self

    def transpose(self):
        _call("u10_tensor_transpose", self.h)
        return self

    def __imul__(self, a):
        _call("u10_tensor_scale", self.h, float(a))
        return self

    def randomize(self, seed: int):
        _call("u10_tensor_randomize", self.h, int(seed))
        return self

    def norm(self) -> float:
        v = ctypes.c_double()
        _call("u10_tensor_norm", self.h, ctypes.byref(v))
        return v.value

    def device_ptr(self) -> int:
        p = ctypes.c_void_p()
        _call("u10_tensor_device_ptr", self.h, ctypes.byref(p))
        return p.value

    def __mul__(self, other):
        if isinstance(other, UniTensor):      # operator*(UniTensor, UniTensor) of the C++ API
            h = ctypes.c_void_p()
            _call("u10_tensor_mul", self.h, other.h, ctypes.byref(h))
            return UniTensor(_handle=h.value)
        t = self.copy()
        t *= other
        return t

    def exSwap(self, other) -> List[tuple]:
        n = ctypes.c_int()
        buf = np.zeros(2 * 256, np.int32)
        _call("u10_tensor_ex_swap", self.h, other.h, buf.ctypes.data_as(P), 256, ctypes.byref(n))
        if n.value > 256:
            buf = np.zeros(2 * n.value, np.int32)
            _call("u10_tensor_ex_swap", self.h, other.h, buf.ctypes.data_as(P), n.value, ctypes.byref(n))
        return [(int(buf[2 * i]), int(buf[2 * i + 1])) for i in range(n.value)]

    def addGate(self, swaps):
        a = np.array([v for sw in swaps for v in sw] or [0], np.int32)
        _call("u10_tensor_add_gate", self.h, len(swaps), a.ctypes.data_as(P))
        return self


def contract(a: UniTensor, b: UniTensor, fast: bool = False) -> UniTensor:
    h = ctypes.c_void_p()
    _call("u10_contract", a.h, b.h, 1 if fast else 0, ctypes.byref(h))
    return UniTensor(_handle=h.value)


def otimes(a: UniTensor, b: UniTensor) -> UniTensor:
    h = ctypes.c_void_p()
    _call("u10_otimes", a.h, b.h, ctypes.byref(h))
    return UniTensor(_handle=h.value)


class Network:
    def __init__(self, fname: str):
        h = ctypes.c_void_p()
        _call("u10_network_create", fname.encode(), ctypes.byref(h))
        self.h = h.value

    def putTensor(self, name: str, T: UniTensor):
        _call("u10_network_put", self.h, name.encode(), T.h, 0)

    def putTensorT(self, name: str, T: UniTensor):
        _call("u10_network_put", self.h, name.encode(), T.h, 1)

    def launch(self, name: str = "") -> UniTensor:
        h = ctypes.c_void_p()
        _call("u10_network_launch", self.h, name.encode(), ctypes.byref(h))
        return UniTensor(_handle=h.value)

    def setGraph(self, on: bool = True):
        _call("u10_network_set_graph", self.h, 1 if on else 0)

    def tree(self) -> str:
        buf = ctypes.create_string_buffer(4096)
        _call("u10_network_tree", self.h, buf, 4096)
        return buf.value.decode()

    def __str__(self) -> str:
        """operator<< of the C++ API: the reference's network print-out."""
        buf = ctypes.create_string_buffer(1 << 16)
        _call("u10_network_print", self.h, buf, 1 << 16)
        return buf.value.decode()

    def __del__(self):
        try:
            if getattr(self, "h", None):
                load().u10_network_free(self.h)
                self.h = None
        except Exception:
            pass


def rdotr(a, a_shape, a_diag, b, b_shape, b_diag):
    """RDotR of the C++ API on host arrays; a diagonal operand is its vector
    of min(rows, cols) entries.  Returns (array, shape, diag)."""
    a, pa = _arr(a, np.float64)
    b, pb = _arr(b, np.float64)
    cap = max(1, int(a_shape[0]) * int(b_shape[1]))
    out = np.zeros(cap)
    cr, cc, cd = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
    _call("u10_rdotr", a_shape[0], a_shape[1], int(a_diag), pa, b_shape[0], b_shape[1], int(b_diag), pb,
          ctypes.byref(cr), ctypes.byref(cc), ctypes.byref(cd), out.ctypes.data_as(P), cap)
    n = min(cr.value, cc.value) if cd.value else cr.value * cc.value
    res = out[:n].copy()
    return (res if cd.value else res.reshape(cr.value, cc.value)), (cr.value, cc.value), bool(cd.value)


def svd(A):
    """A = U diag(S) VT (Block::svd); complex A through the complex SVD."""
    if np.iscomplexobj(A):
        a = np.ascontiguousarray(A, dtype=np.complex128)
        M, N = a.shape
        k = min(M, N)
        U, Sv, VT = np.empty((M, k), np.complex128), np.empty(k), np.empty((k, N), np.complex128)
        _call("u10_svd_c128", _cplx(a).ctypes.data_as(P), M, N, _cplx(U).ctypes.data_as(P),
              Sv.ctypes.data_as(P), _cplx(VT).ctypes.data_as(P))
        return U, Sv, VT
    a, p = _arr(A, np.float64)
    M, N = a.shape
    k = min(M, N)
    U, Sv, VT = np.empty((M, k)), np.empty(k), np.empty((k, N))
    _call("u10_svd", p, M, N, U.ctypes.data_as(P), Sv.ctypes.data_as(P), VT.ctypes.data_as(P))
    return U, Sv, VT


def eigh(A):
    """eigenvalues ascending and eigenvector rows (eigSyDecompose); a complex
    (Hermitian) A gives rows E with A = E^H diag(w) E."""
    if np.iscomplexobj(A):
        a = np.ascontiguousarray(A, dtype=np.complex128)
        n = a.shape[0]
        w, v = np.empty(n), np.empty((n, n), np.complex128)
        _call("u10_eigh_c128", _cplx(a).ctypes.data_as(P), n, w.ctypes.data_as(P), _cplx(v).ctypes.data_as(P))
        return w, v
    a, p = _arr(A, np.float64)
    n = a.shape[0]
    w, v = np.empty(n), np.empty((n, n))
    _call("u10_eigh", p, n, w.ctypes.data_as(P), v.ctypes.data_as(P))
    return w, v


def lanczosEigh(A, psi, max_iter: int = 1000, err_tol: float = 5e-15):
    """Lowest eigenpair of a square matrix by Lanczos (Block.cpp:474-506):
    returns (E0, ground state, steps taken); psi is the start vector."""
    a, p = _arr(A, np.float64)
    if a.ndim != 2 or a.shape[0] != a.shape[1]:
        raise Uni10Error("lanczosEigh: Cannot perform Lanczos algorithm to find the lowest eigen value and "
                         "eigen vector on a non-square matrix.")
    n = a.shape[0]
    v = np.array(psi, dtype=np.float64).reshape(-1).copy()
    if v.size != n:
        raise ValueError("Error in Lanczos initial vector psi. The vector dimension does not match with the "
                         "number of the columns.")
    e0 = ctypes.c_double()
    it = ctypes.c_uint64()
    _call("u10_lanczos_eigh", p, n, v.ctypes.data_as(P), int(max_iter), float(err_tol), ctypes.byref(e0),
          ctypes.byref(it))
    return e0.value, v, int(it.value)


def exph(x: float, H):
    a, p = _arr(H, np.float64)
    n = a.shape[0]
    out = np.empty((n, n))
    _call("u10_exph", float(x), p, n, out.ctypes.data_as(P))
    return out


def u1_bond_sectors(chi: int, d: int, bond: int):
    """(charges, dims) the U(1) chains of BatchTebd give bond `bond` (0:
    lambda_A, 1: lambda_B) at bond dimension chi (host-side rule)."""
    n = ctypes.c_int(0)
    _call("u10_u1_bond_sectors", int(chi), int(d), int(bond), None, None, ctypes.byref(n))
    q = np.empty(n.value, np.int32)
    m = np.empty(n.value, np.int32)
    _call("u10_u1_bond_sectors", int(chi), int(d), int(bond), q.ctypes.data_as(P), m.ctypes.data_as(P),
          ctypes.byref(n))
    return q, m


class BatchTebd:
    """Many independent XXZ chains advanced together: every stage of an
    update is one launch for all chains (grouped GEMMs, batched SVD).
    d[i] = 2 (spin 1/2) or 3 (spin 1), sym[i] = 0 (dense) or 1 (U(1));
    None = all 2 / all dense."""

    def __init__(self, chi, delta, seed, tau: float, netdir: Optional[str] = None, gemm_mode: int = 0, d=None,
                 sym=None):
        chi = np.ascontiguousarray(chi, dtype=np.int32)
        delta = np.ascontiguousarray(delta, dtype=np.float64)
        seed = np.ascontiguousarray(seed, dtype=np.uint64)
        dd = None if d is None else np.ascontiguousarray(d, dtype=np.int32)
        sy = None if sym is None else np.ascontiguousarray(sym, dtype=np.int32)
        h = ctypes.c_void_p()
        _call("u10_batch_tebd_create_ex", len(chi), chi.ctypes.data_as(P), delta.ctypes.data_as(P),
              seed.ctypes.data_as(P), None if dd is None else dd.ctypes.data_as(P),
              None if sy is None else sy.ctypes.data_as(P), float(tau), int(gemm_mode), (netdir or "").encode(),
              ctypes.byref(h))
        self.h = h.value
        self.chi = chi
        self.d = np.full(len(chi), 2, np.int32) if dd is None else dd
        self.sym = np.zeros(len(chi), np.int32) if sy is None else sy

    def sectors(self, chain: int, bond: int):
        """(charges, dims) of a U(1) chain's bond (0: lambda_A, 1: lambda_B);
        empty arrays for a dense chain."""
        n = ctypes.c_int(0)
        _call("u10_batch_tebd_sectors", self.h, int(chain), int(bond), None, None, ctypes.byref(n))
        q = np.empty(n.value, np.int32)
        m = np.empty(n.value, np.int32)
        _call("u10_batch_tebd_sectors", self.h, int(chain), int(bond), q.ctypes.data_as(P),
              m.ctypes.data_as(P), ctypes.byref(n))
        return q, m

    def step(self, n: int = 1):
        _call("u10_batch_tebd_step", self.h, int(n))

    def launch(self, n: int = 1):
        """Enqueue n steps on the engine's own stream without waiting."""
        _call("u10_batch_tebd_launch", self.h, int(n))

    def wait(self):
        _call("u10_batch_tebd_wait", self.h)

    def ready(self) -> bool:
        """True once the launched steps have finished (a poll; never blocks)."""
        out = ctypes.c_int(0)
        _call("u10_batch_tebd_ready", self.h, ctypes.byref(out))
        return bool(out.value)

    def set_graph(self, on: bool = True):
        _call("u10_batch_tebd_set_graph", self.h, int(bool(on)))

    def set_stream(self, stream):
        """Run the engine's work on `stream` (None: its own stream)."""
        _call("u10_batch_tebd_set_stream", self.h, stream)

    def set_active(self, chain: int, on: bool = True):
        """An inactive chain is skipped on the device (its state is kept)."""
        _call("u10_batch_tebd_set_active", self.h, int(chain), int(bool(on)))

    def load(self, chain: int, delta: float, seed: int):
        _call("u10_batch_tebd_load", self.h, int(chain), float(delta), int(seed))

    def energy(self, chain: int) -> float:
        v = ctypes.c_double()
        _call("u10_batch_tebd_energy", self.h, int(chain), ctypes.byref(v))
        return v.value

    def lam(self, chain: int, which: int) -> np.ndarray:
        n = ctypes.c_int()
        _call("u10_batch_tebd_lambda", self.h, int(chain), int(which), None, ctypes.byref(n))
        out = np.empty(n.value)
        _call("u10_batch_tebd_lambda", self.h, int(chain), int(which), out.ctypes.data_as(P), ctypes.byref(n))
        return out

    def trunc_err(self, chain: int) -> float:
        v = ctypes.c_double()
        _call("u10_batch_tebd_trunc_err", self.h, int(chain), ctypes.byref(v))
        return v.value

    def info(self):
        f, s = ctypes.c_double(), ctypes.c_int()
        _call("u10_batch_tebd_info", self.h, ctypes.byref(f), ctypes.byref(s))
        return {"gemm_flops_per_step": f.value, "last_sweeps": s.value}

    def __del__(self):
        try:
            if getattr(self, "h", None):
                load().u10_batch_tebd_free(self.h)
                self.h = None
        except Exception:
            pass


class ITebd:
    """The reference example's iTEBD (Networks + SVD truncation) in C++."""

    def __init__(self, chi: int, d: int, delta: float, tau: float, seed: int, netdir: Optional[str] = None):
        h = ctypes.c_void_p()
        _call("u10_itebd_create", int(chi), int(d), float(delta), float(tau), int(seed),
              (netdir or "").encode(), ctypes.byref(h))
        self.h = h.value
        self.chi = chi

    def step(self, n: int = 1):
        _call("u10_itebd_step", self.h, int(n))

    def set_graph(self, on: bool = True):
        _call("u10_itebd_set_graph", self.h, 1 if on else 0)

    def energy(self) -> float:
        v = ctypes.c_double()
        _call("u10_itebd_energy", self.h, ctypes.byref(v))
        return v.value

    def lam(self, which: int) -> np.ndarray:
        n = ctypes.c_int()
        _call("u10_itebd_lambda", self.h, which, None, ctypes.byref(n))
        out = np.empty(n.value)
        _call("u10_itebd_lambda", self.h, which, out.ctypes.data_as(P), ctypes.byref(n))
        return out

    def trunc_err(self) -> float:
        v = ctypes.c_double()
        _call("u10_itebd_trunc_err", self.h, ctypes.byref(v))
        return v.value

    def __del__(self):
        try:
            if getattr(self, "h", None):
                load().u10_itebd_free(self.h)
                self.h = None
        except Exception:
            pass
