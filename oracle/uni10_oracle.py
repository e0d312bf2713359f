"""CPU oracle for the Uni10 tensor-contraction hot path.

TEST INFRASTRUCTURE ONLY.  This package is the checker: the tests, the
smoke check and the bench's ``cpu_baseline`` leg import it; the product
(``trim_amd`` and its HIP library) never does.

It restates, in Python/numpy (and the C loops of ``uni10_cpu.c``), the
reference algorithms of Uni10 v1.0 that the hot path runs through.  Every
function cites the reference lines it follows (paths relative to the
reference checkout, ``S/`` = ``uni10/src/uni10/``).

Floating-point kernels call the same LAPACK/BLAS routines the reference calls
(``dgemm`` through numpy's matmul, ``dgesvd`` and ``dsyev`` through SciPy);
integer/index work (block layout, permutes) is restated exactly.

Pinning (what each part is checked against; see tests/test_oracle.py):
  * Qnum / Bond           - reference unit-test known answers (testQnum.cpp,
                            testBond.cpp).
  * grouping/setRawElem   - golden tensor dumps egU1_H_U1 (rank-4 U(1)),
                            egU3_W / egU3_WT (rank-3 U(1), transposed pair).
  * dense permute         - numpy.transpose (textbook routine) and the
                            _reshapeElem formula (two independent statements).
  * symmetric permute     - densify-and-compare: raw(permute(T)) ==
                            transpose(raw(T)), plus the egU3 transpose pair.
  * contract              - dense tensordot on densified tensors.
  * svd                   - golden mat_U/mat_Sigma/mat_VT dump: conventions,
                            reconstruction, orthonormality, ordering.
  * eigh / exph           - closed form of the two-site spin-1/2 Heisenberg
                            gate (eigenvalues 1/4 x3, -3/4).
  * lanczos_ev            - the lowest eigenvalue from numpy.linalg.eigh
                            (textbook routine) and the closed form -2 of the
                            four-site Heisenberg ring; the eigenvector by its
                            residual.
  * Network               - Simple.net and egN1_network known answers; ORDER
                            brackets vs greedy plan on the same tensors; the
                            greedy tree, node element counts and label order
                            printed in Network.h:145-193 (print_tree).
  * iTEBD energy          - Bethe-ansatz bound 1/4 - ln 2 and the two-site
                            exact ground energy -3/4 at chi=1.
  * rec_swap              - inversion-set identity and permutation sign.
  * swap_signs_setup      - the convention (row bonds, then column bonds in
                            reverse, reordered as Grassmann generators;
                            UniTensorReal.cpp:657-681) against an independent
                            per-element reordering sign (cycle decomposition),
                            round trips and composition across row-bond
                            counts (bit-exact), and hand-derived rank-2 cases.
  * add_gate / add_swap   - addGate against the per-element parity product;
                            addSwap's gate list on a three-tensor chain (hand
                            derived) and order independence of the launched
                            fermionic result (the Network.h network under
                            three trees with different gates; a chain against
                            the plain matrix product).
"""
from __future__ import annotations

import ctypes
import itertools
import math
import os
import subprocess
from collections import OrderedDict, defaultdict
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

BD_IN = 1
BD_OUT = -1
PRT_EVEN, PRT_ODD = 0, 1
PRTF_EVEN, PRTF_ODD = 0, 1
U1_UPB, U1_LOB = 1000, -1000

_HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------------------
# C helpers (uni10_cpu.c), built on first use with gcc.
# ---------------------------------------------------------------------------
_LIB = None


def build_c(force: bool = False) -> str:
    """Compile uni10_cpu.c to liboracle_cpu.so next to it; return the path."""
    src = os.path.join(_HERE, "uni10_cpu.c")
    out = os.path.join(_HERE, "liboracle_cpu.so")
    if force or not os.path.exists(out) or os.path.getmtime(out) < os.path.getmtime(src):
        tmp = out + ".tmp%d" % os.getpid()
        subprocess.check_call(["gcc", "-O2", "-fPIC", "-shared", "-o", tmp, src])
        os.replace(tmp, out)
    return out


def lib():
    global _LIB
    if _LIB is None:
        L = ctypes.CDLL(build_c())
        P = ctypes.c_void_p
        L.ora_permute_dense.argtypes = [P, P, ctypes.c_int, P, P]
        L.ora_reshape_elem.argtypes = [P, P, ctypes.c_int, P, P]
        L.ora_permute_moves.argtypes = [P, P, ctypes.c_int, ctypes.c_int64, P, P, P, P]
        L.ora_gemm_rowmajor.argtypes = [P, P, P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]
        _LIB = L
    return _LIB


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------------------
# Qnum  (S/datatype/lib/Qnum.cpp:33-100)
# ---------------------------------------------------------------------------
class Qnum:
    """U(1) charge, Z2 parity and fermion parity; ordered by hash = 4*U1+2*prt+prtF.

    Follows Qnum.cpp:34-59 (range check), :62-67 (ordering by hash), :71-74
    (negation of U1 only), :76-79 (fusion: add U1, XOR parities), :100 (hash).
    Constructing a fermion-odd Qnum sets the process-wide fermionic flag
    (Qnum.cpp:33,54-55) -- restated as ``Qnum.fermionic``.
    """

    __slots__ = ("U1", "prt", "prtF")
    fermionic = False

    def __init__(self, U1: int = 0, prt: int = PRT_EVEN, prtF: int = PRTF_EVEN):
        if not (U1_LOB < U1 < U1_UPB):
            raise ValueError("U1 is out of range. %d < U1 < %d." % (U1_LOB, U1_UPB))
        self.U1, self.prt, self.prtF = int(U1), int(prt), int(prtF)
        if self.prtF == PRTF_ODD:
            Qnum.fermionic = True

    def hash(self) -> int:
        return self.U1 * 4 + self.prt * 2 + self.prtF

    def key(self):
        return self.hash()

    def __lt__(self, o):
        return self.hash() < o.hash()

    def __le__(self, o):
        return self.hash() <= o.hash()

    def __eq__(self, o):
        return isinstance(o, Qnum) and (self.U1, self.prt, self.prtF) == (o.U1, o.prt, o.prtF)

    def __hash__(self):
        return self.hash()

    def __neg__(self):
        return Qnum(-self.U1, self.prt, self.prtF)

    def __mul__(self, o):
        return Qnum(self.U1 + o.U1, self.prt ^ o.prt, self.prtF ^ o.prtF)

    def triple(self):
        return (self.U1, self.prt, self.prtF)

    def __repr__(self):
        return "Q(%d,%d,%d)" % (self.U1, self.prt, self.prtF)


Q0 = Qnum(0)


# ---------------------------------------------------------------------------
# Bond  (S/data-structure/lib/Bond.cpp:90-252)
# ---------------------------------------------------------------------------
class Bond:
    """Run-length list of Qnums; only adjacent equal Qnums merge (Bond.cpp:90-110)."""

    def __init__(self, tp: int, qnums_or_dim):
        self.type = tp
        if isinstance(qnums_or_dim, (int, np.integer)):
            qnums = [Qnum(0)] * int(qnums_or_dim)
        else:
            qnums = list(qnums_or_dim)
        if len(qnums) == 0:
            raise ValueError("Cannot create a bond of dimension 0.")
        self.Qnums: List[Qnum] = []
        self.Qdegs: List[int] = []
        self.offsets: List[int] = []
        self.dim = 0
        for i, q in enumerate(qnums):
            if i == 0 or not (q == qnums[i - 1]):
                self.Qnums.append(q)
                self.Qdegs.append(1)
                self.offsets.append(self.dim)
            else:
                self.Qdegs[-1] += 1
            self.dim += 1

    @classmethod
    def from_runs(cls, tp, qnums, degs):
        full = []
        for q, d in zip(qnums, degs):
            full.extend([q] * d)
        return cls(tp, full)

    def copy(self) -> "Bond":
        b = Bond.__new__(Bond)
        b.type, b.dim = self.type, self.dim
        b.Qnums, b.Qdegs, b.offsets = list(self.Qnums), list(self.Qdegs), list(self.offsets)
        return b

    def change(self, tp: int) -> "Bond":
        """Bond.cpp:150-157: flipping in<->out negates every charge."""
        if self.type != tp:
            self.Qnums = [-q for q in self.Qnums]
            self.type = tp
        return self

    def Qlist(self) -> List[Qnum]:
        out = []
        for q, d in zip(self.Qnums, self.Qdegs):
            out.extend([q] * d)
        return out

    def degeneracy(self) -> "OrderedDict[Qnum, int]":
        """Bond.cpp:125-134, returned in Qnum order."""
        h: Dict[Qnum, int] = {}
        for q, d in zip(self.Qnums, self.Qdegs):
            h[q] = h.get(q, 0) + d
        return OrderedDict(sorted(h.items(), key=lambda kv: kv[0].hash()))

    def combine(self, bd: "Bond") -> "Bond":
        """Bond.cpp:163-196: fuse every state of self with every run of bd."""
        bd = bd.copy().change(self.type)
        qnums: List[Qnum] = []
        qdegs: List[int] = []
        offsets: List[int] = []
        dim = 0
        for q, dq in zip(self.Qnums, self.Qdegs):
            for _ in range(dq):
                for qq, dqq in zip(bd.Qnums, bd.Qdegs):
                    fused = q * qq
                    if not qnums or not (fused == qnums[-1]):
                        qnums.append(fused)
                        qdegs.append(dqq)
                        offsets.append(dim)
                    else:
                        qdegs[-1] += dqq
                    dim += dqq
        self.Qnums, self.Qdegs, self.offsets, self.dim = qnums, qdegs, offsets, dim
        return self

    def __eq__(self, o):
        return self.type == o.type and self.Qnums == o.Qnums and self.Qdegs == o.Qdegs

    def __repr__(self):
        return "Bond(%s, %s)" % ("IN" if self.type == BD_IN else "OUT",
                                 list(zip(self.Qnums, self.Qdegs)))


def combine_bonds(tp: int, bds: Sequence[Bond]) -> Bond:
    """Bond.cpp:198-238: right-to-left pairwise combine with alternating buffers."""
    if len(bds) == 0:
        raise ValueError("There should be at least one bond in the input vector to be combined.")
    if len(bds) == 1:
        return bds[0].copy().change(tp)
    n = len(bds)
    out1 = bds[n - 1].copy()
    out2 = bds[n - 2].copy()
    out2.change(tp)
    out2.combine(out1)
    b = 0
    for b in range(n - 2):
        if b % 2 == 0:
            out1 = bds[n - 3 - b].copy()
            out1.change(tp)
            out1.combine(out2)
        else:
            out2 = bds[n - 3 - b].copy()
            out2.change(tp)
            out2.combine(out1)
    else:
        b = n - 2
    return out2 if b % 2 == 0 else out1


# ---------------------------------------------------------------------------
# Block layout  (S/tensor-network/lib/UniTensorReal.cpp:1169-1321)
# ---------------------------------------------------------------------------
class Layout:
    """Block plan of a tensor with the given bonds.

    Restates ``UniTensor::grouping``: row combos are an odometer over the
    run-length Qnum indices of the in-bonds (last bond fastest,
    UniTensorReal.cpp:1203-1234); column combos likewise over the out-bonds,
    kept only where the row charge exists (:1244-1277); blocks are the charges
    present on both sides, in Qnum hash order, concatenated in one buffer
    (:1286-1310, initBlocks :1365-1374); QidxEnc enumerates the
    (row combo, column combo) pairs of equal charge in increasing
    qidx = RQidx*CQdim + CQidx (:1311-1319).
    """

    def __init__(self, bonds: Sequence[Bond]):
        self.bonds = [b.copy() for b in bonds]
        nb = len(self.bonds)
        self.blocks: "OrderedDict[Qnum, Tuple[int, int, int]]" = OrderedDict()  # q -> (rows, cols, off)
        self.RQidx2Blk: Dict[int, Qnum] = {}
        self.RQidx2Off: Dict[int, int] = {}
        self.RQidx2Dim: Dict[int, int] = {}
        self.CQidx2Off: Dict[int, int] = {}
        self.CQidx2Dim: Dict[int, int] = {}
        self.QidxEnc: "OrderedDict[int, int]" = OrderedDict()
        if nb == 0:
            self.RBondNum, self.RQdim, self.CQdim = 0, 0, 0
            self.blocks[Qnum(0)] = (1, 1, 0)
            self.elemNum = 1
            return
        nrow = 0
        seen_out = False
        self.RQdim = self.CQdim = 1
        for b in self.bonds:
            if b.type == BD_IN:
                if seen_out:
                    raise ValueError("BD_OUT bonds must be placed after all BD_IN bonds.")
                self.RQdim *= len(b.Qnums)
                nrow += 1
            else:
                self.CQdim *= len(b.Qnums)
                seen_out = True
        self.RBondNum = nrow
        rbonds, cbonds = self.bonds[:nrow], self.bonds[nrow:]

        row_mdim: Dict[Qnum, int] = {}
        row_q2idx: Dict[Qnum, List[int]] = defaultdict(list)
        tmpR_dim = [1] * self.RQdim
        tmpR_off = [0] * self.RQdim
        if nrow:
            for ridx, combo in enumerate(itertools.product(*[range(len(b.Qnums)) for b in rbonds])):
                q, dim = Qnum(0), 1
                for b, qi in zip(rbonds, combo):
                    q = q * b.Qnums[qi]
                    dim *= b.Qdegs[qi]
                tmpR_off[ridx] = row_mdim.get(q, 0)
                tmpR_dim[ridx] = dim
                row_mdim[q] = tmpR_off[ridx] + dim
                row_q2idx[q].append(ridx)
        else:
            row_mdim[Qnum(0)] = 1
            row_q2idx[Qnum(0)].append(0)

        col_mdim: Dict[Qnum, int] = {}
        col_q2idx: Dict[Qnum, List[int]] = defaultdict(list)
        tmpC_dim = [1] * self.CQdim
        tmpC_off = [0] * self.CQdim
        if len(cbonds):
            for cidx, combo in enumerate(itertools.product(*[range(len(b.Qnums)) for b in cbonds])):
                q, dim = Qnum(0), 1
                for b, qi in zip(cbonds, combo):
                    q = q * b.Qnums[qi]
                    dim *= b.Qdegs[qi]
                if q in row_mdim:
                    tmpC_off[cidx] = col_mdim.get(q, 0)
                    tmpC_dim[cidx] = dim
                    col_mdim[q] = tmpC_off[cidx] + dim
                    col_q2idx[q].append(cidx)
        else:
            if Qnum(0) in row_mdim:
                col_mdim[Qnum(0)] = 1
                col_q2idx[Qnum(0)].append(0)

        qidx_set = set()
        off = 0
        for q in sorted(col_mdim.keys(), key=lambda x: x.hash()):
            rows, cols = row_mdim[q], col_mdim[q]
            self.blocks[q] = (rows, cols, off)
            off += rows * cols
            for ridx in row_q2idx[q]:
                self.RQidx2Blk[ridx] = q
                for cidx in col_q2idx[q]:
                    self.RQidx2Dim[ridx] = tmpR_dim[ridx]
                    self.RQidx2Off[ridx] = tmpR_off[ridx]
                    self.CQidx2Dim[cidx] = tmpC_dim[cidx]
                    self.CQidx2Off[cidx] = tmpC_off[cidx]
                    qidx_set.add(ridx * self.CQdim + cidx)
        enc = 0
        for ridx in sorted(self.RQidx2Dim):
            for cidx in sorted(self.CQidx2Dim):
                qidx = ridx * self.CQdim + cidx
                if qidx in qidx_set:
                    self.QidxEnc[qidx] = enc
                    enc += self.RQidx2Dim[ridx] * self.CQidx2Dim[cidx]
        self.elemNum = off
        if not self.blocks:
            raise ValueError("There is no symmetry block with the given bonds")

    # helpers --------------------------------------------------------------
    def qidx_decode(self, qidx: int) -> List[int]:
        idxs = [0] * len(self.bonds)
        tmp = qidx
        for b in range(len(self.bonds) - 1, -1, -1):
            n = len(self.bonds[b].Qnums)
            idxs[b] = tmp % n
            tmp //= n
        return idxs

    def subblock(self, qidx: int):
        """(element offset of the sub-block corner, block cols, sub rows, sub cols)."""
        ridx, cidx = divmod(qidx, self.CQdim)
        q = self.RQidx2Blk[ridx]
        rows, cols, boff = self.blocks[q]
        eoff = boff + self.RQidx2Off[ridx] * cols + self.CQidx2Off[cidx]
        return eoff, cols, self.RQidx2Dim[ridx], self.CQidx2Dim[cidx]


# ---------------------------------------------------------------------------
# UniTensor (real, FP64)  (S/tensor-network/lib/UniTensorReal.cpp)
# ---------------------------------------------------------------------------
class UniTensor:
    """Real block-sparse tensor: bonds, labels, blocks over one buffer."""

    def __init__(self, bonds: Sequence[Bond] = (), name: str = "", elem=None, scalar=None):
        self.layout = Layout(bonds)
        self.bonds = self.layout.bonds
        self.labels = list(range(len(self.bonds)))
        self.name = name
        self.elem = np.zeros(self.layout.elemNum, dtype=np.float64)
        self.have_elem = False
        if elem is not None:
            self.set_elem(elem)
        if scalar is not None:
            self.elem[0] = scalar
            self.have_elem = True

    # basic accessors
    @property
    def RBondNum(self):
        return self.layout.RBondNum

    def bond_num(self):
        return len(self.bonds)

    def in_bond_num(self):
        return self.layout.RBondNum

    def elem_num(self):
        return self.layout.elemNum

    def copy(self) -> "UniTensor":
        t = UniTensor.__new__(UniTensor)
        t.layout = self.layout
        t.bonds = self.layout.bonds
        t.labels = list(self.labels)
        t.name = self.name
        t.elem = self.elem.copy()
        t.have_elem = self.have_elem
        return t

    def set_label(self, labels):
        if len(labels) != len(self.bonds):
            raise ValueError("The size of input vector(labels) does not match for the number of bonds.")
        if len(set(labels)) != len(labels):
            raise ValueError("Labels must be distinct.")
        self.labels = list(labels)
        return self

    def set_elem(self, e):
        e = np.asarray(e, dtype=np.float64).reshape(-1)
        if e.size != self.layout.elemNum:
            raise ValueError("element count mismatch")
        self.elem[:] = e
        self.have_elem = True

    def block_qnums(self):
        return list(self.layout.blocks.keys())

    def get_block(self, q: Optional[Qnum] = None) -> np.ndarray:
        q = Qnum(0) if q is None else q
        if q not in self.layout.blocks:
            raise KeyError("There is no block with the given quantum number %r" % (q,))
        r, c, off = self.layout.blocks[q]
        return self.elem[off:off + r * c].reshape(r, c).copy()

    def put_block(self, mat, q: Optional[Qnum] = None):
        """UniTensorReal.cpp:206-239 (dense matrix case)."""
        q = Qnum(0) if q is None else q
        if q not in self.layout.blocks:
            raise KeyError("There is no block with the given quantum number %r" % (q,))
        r, c, off = self.layout.blocks[q]
        mat = np.asarray(mat, dtype=np.float64)
        if mat.shape != (r, c):
            raise ValueError("The dimension of input matrix does not match for the dimension of the block")
        self.elem[off:off + r * c] = mat.reshape(-1)
        self.have_elem = True

    # raw (dense) element access ------------------------------------------
    def set_raw_elem(self, raw):
        """UniTensorReal.cpp:96-174: scatter a dense row-major tensor into the blocks.

        Each QidxEnc entry is a sub-block: its dense corner is
        sum_b rAcc[b]*offsets[Q_idxs[b]]; its extents are the Qdegs; its
        elements land row-major at the sub-block position inside the block.
        """
        L = self.layout
        dims = [b.dim for b in self.bonds]
        raw = np.asarray(raw, dtype=np.float64).reshape(dims)
        for qidx in L.QidxEnc:
            qi = L.qidx_decode(qidx)
            sl = tuple(slice(b.offsets[i], b.offsets[i] + b.Qdegs[i]) for b, i in zip(self.bonds, qi))
            eoff, bcols, sr, sc = L.subblock(qidx)
            sub = raw[sl].reshape(sr, sc)
            for r in range(sr):
                self.elem[eoff + r * bcols: eoff + r * bcols + sc] = sub[r]
        self.have_elem = True

    def at(self, idxs: Sequence[int]) -> float:
        """UniTensorReal.cpp:1098-1165: element of the dense tensor, 0 outside the blocks."""
        L = self.layout
        if len(idxs) != len(self.bonds):
            raise ValueError("The size of input indices array does not match with the number of the bonds.")
        qidxs = []
        for b, i in zip(self.bonds, idxs):
            if not (0 <= i < b.dim):
                raise IndexError("The input indices are out of range.")
            q = max(k for k, o in enumerate(b.offsets) if o <= i)
            qidxs.append(q)
        qoff = 0
        for b, q in zip(self.bonds, qidxs):
            qoff = qoff * len(b.Qnums) + q
        if qoff not in L.QidxEnc:
            return 0.0
        eoff, bcols, sr, sc = L.subblock(qoff)
        cnt = 0
        for b, q, i in zip(self.bonds, qidxs, idxs):
            cnt = cnt * b.Qdegs[q] + (i - b.offsets[q])
        return float(self.elem[eoff + (cnt // sc) * bcols + cnt % sc])

    def get_raw_elem(self) -> np.ndarray:
        """UniTensorReal.cpp:241-280: dense (rows x cols) matrix of the tensor."""
        if not self.bonds:
            return self.elem.reshape(1, 1).copy()
        L = self.layout
        dims = [b.dim for b in self.bonds]
        raw = np.zeros(dims, dtype=np.float64)
        for qidx in L.QidxEnc:
            qi = L.qidx_decode(qidx)
            sl = tuple(slice(b.offsets[i], b.offsets[i] + b.Qdegs[i]) for b, i in zip(self.bonds, qi))
            eoff, bcols, sr, sc = L.subblock(qidx)
            sub = np.empty((sr, sc))
            for r in range(sr):
                sub[r] = self.elem[eoff + r * bcols: eoff + r * bcols + sc]
            raw[sl] = sub.reshape([b.Qdegs[i] for b, i in zip(self.bonds, qi)])
        rows = int(np.prod([b.dim for b in self.bonds if b.type == BD_IN])) if self.bonds else 1
        return raw.reshape(rows, -1)

    # permute ----------------------------------------------------------------
    def permute(self, new_labels: Sequence[int], row_bond_num: int) -> "UniTensor":
        """UniTensorReal.cpp:532-768 (in place, like the reference)."""
        if not self.bonds:
            raise ValueError("There is no bond in the tensor(scalar) to permute.")
        nb = len(self.bonds)
        if len(new_labels) != nb:
            raise ValueError("The size of the input new labels does not match for the number of bonds.")
        rsp = [0] * nb
        cnt = 0
        for i in range(nb):
            for j in range(nb):
                if self.labels[i] == new_labels[j]:
                    rsp[j] = i
                    cnt += 1
        if cnt != nb:
            raise ValueError("The input new labels do not 1-1 correspond to the labels of the tensor.")
        inorder = all(rsp[i] == i for i in range(1, nb))
        if inorder and self.RBondNum == row_bond_num:
            return self
        out_bonds = [self.bonds[rsp[b]].copy() for b in range(nb)]
        without_sym = all(len(b.Qnums) == 1 for b in self.bonds)
        for b in range(nb):
            out_bonds[b].change(BD_IN if b < row_bond_num else BD_OUT)
        out = UniTensor(out_bonds, self.name)
        if self.have_elem:
            if without_sym:
                if not inorder:
                    dims = np.array([b.dim for b in self.bonds], dtype=np.int64)
                    perm = np.array(rsp, dtype=np.int32)
                    src = np.ascontiguousarray(self.elem)
                    lib().ora_permute_dense(_ptr(src), _ptr(out.elem), nb, _ptr(dims), _ptr(perm))
                else:
                    out.elem[:] = self.elem
            else:
                moves = permute_moves(self, out, rsp, new_labels)
                run_moves(self.elem, out.elem, nb, moves)
            out.have_elem = True
        self.layout, self.bonds, self.elem, self.have_elem = out.layout, out.bonds, out.elem, out.have_elem
        self.labels = list(new_labels)
        return self

    def permute_row(self, row_bond_num: int) -> "UniTensor":
        return self.permute(list(self.labels), row_bond_num)

    def transpose(self) -> "UniTensor":
        """UniTensorReal.cpp:448-506: swap the row and column bond groups, transpose each block.

        Bond types are overwritten without negating charges (:475-480).
        """
        if not self.bonds:
            raise ValueError("There is no bond in the tensor(scalar) to perform transposition.")
        nb = len(self.bonds)
        rb = 0
        for b in self.bonds:
            if b.type == BD_IN:
                rb += 1
            else:
                break
        cb = nb - rb
        rsp = [rb + b if b < cb else b - cb for b in range(nb)]
        out_bonds, out_labels = [], []
        for b in range(nb):
            nbd = self.bonds[rsp[b]].copy()
            nbd.type = BD_IN if b < cb else BD_OUT
            out_bonds.append(nbd)
            out_labels.append(self.labels[rsp[b]])
        out = UniTensor(out_bonds, self.name)
        out.labels = out_labels
        if self.have_elem:
            for q, (r, c, off) in self.layout.blocks.items():
                r2, c2, off2 = out.layout.blocks[q]
                out.elem[off2:off2 + r * c] = self.elem[off:off + r * c].reshape(r, c).T.reshape(-1)
            out.have_elem = True
        self.layout, self.bonds, self.elem, self.labels = out.layout, out.bonds, out.elem, out.labels
        self.have_elem = out.have_elem
        return self

    # arithmetic ---------------------------------------------------------------
    def scale(self, a: float) -> "UniTensor":
        """operator*= (UniTensor.cpp: vectorScal over elem)."""
        self.elem *= a
        return self

    def norm(self) -> float:
        """UniTensorReal.cpp:1388-1396 -> vectorNorm (uni10_lapack_cpu.cpp:397-414)."""
        return float(np.sqrt(np.sum(self.elem * self.elem)))

    def __getitem__(self, i):
        return float(self.elem[i])

    # fermionic swap gates -----------------------------------------------------
    def ex_swap(self, Tb: "UniTensor"):
        """UniTensor::exSwap (UniTensor.cpp:1237-1268): every (shared bond,
        unshared bond) pair of this tensor with respect to Tb."""
        if not (self.bonds and Tb.bonds):
            return []
        shared = [a for a in range(len(self.labels)) if self.labels[a] in Tb.labels]
        left = [a for a in range(len(self.labels)) if self.labels[a] not in Tb.labels]
        return [(i, j) for i in shared for j in left]

    def add_gate(self, swaps) -> "UniTensor":
        """UniTensor::addGate (UniTensorReal.cpp:868-921): each Qidx sub-block
        is multiplied by (-1)^(XOR over gates of prtF(b1) & prtF(b2))."""
        if not self.bonds:
            raise ValueError("Adding swap gates to a tensor without bonds(scalar).")
        if not self.have_elem:
            raise ValueError("Cannot add swap gates to a tensor before setting its elements.")
        L = self.layout
        for qidx in L.QidxEnc:
            q = L.qidx_decode(qidx)
            bit = 0
            for b1, b2 in swaps:
                bit ^= self.bonds[b1].Qnums[q[b1]].prtF & self.bonds[b2].Qnums[q[b2]].prtF
            if bit:
                eoff, bcols, sr, sc = L.subblock(qidx)
                for r in range(sr):
                    self.elem[eoff + r * bcols:eoff + r * bcols + sc] *= -1.0
        return self


def swap_signs_setup(T: UniTensor, out: UniTensor, rsp, new_labels):
    """Fermionic swap list (UniTensorReal.cpp:657-681 -> recSwap, uni10_tools.cpp:52-69)."""
    nb = len(T.bonds)
    RB, oRB = T.RBondNum, out.RBondNum
    in_l = [0] * nb
    out_l = [0] * nb
    ordF = [0] * nb
    for b in range(RB):
        in_l[b] = T.labels[b]
        ordF[b] = b
    for b in range(oRB):
        out_l[b] = new_labels[b]
    for b in range(nb - 1, RB - 1, -1):
        ordF[b] = nb - b + RB - 1
        in_l[ordF[b]] = T.labels[b]
    for b in range(nb - 1, oRB - 1, -1):
        out_l[nb - b + oRB - 1] = new_labels[b]
    rspF = [0] * nb
    for i in range(nb):
        for j in range(nb):
            if in_l[i] == out_l[j]:
                rspF[j] = i
    return rec_swap(rspF, ordF)


def rec_swap(order, ordF):
    """uni10_tools.cpp:52-69: bubble sort, recording each swapped bond pair."""
    ordv = list(order)
    n = len(ordv)
    swaps = []
    for i in range(n - 1):
        for j in range(n - i - 1):
            if ordv[j] > ordv[j + 1]:
                swaps.append((ordF[ordv[j + 1]], ordF[ordv[j]]))
                ordv[j], ordv[j + 1] = ordv[j + 1], ordv[j]
    return swaps


def permute_moves(T: UniTensor, out: UniTensor, rsp, new_labels):
    """Per-Qidx metadata of the symmetric permute (UniTensorReal.cpp:683-738)."""
    L, Lo = T.layout, out.layout
    nb = len(T.bonds)
    swaps = swap_signs_setup(T, out, rsp, new_labels) if Qnum.fermionic else []
    qot_acc = [1] * nb
    for b in range(nb - 1, 0, -1):
        qot_acc[b - 1] = qot_acc[b] * len(out.bonds[b].Qnums)
    meta, subd, acc, sign = [], [], [], []
    for qidx in L.QidxEnc:
        qin = L.qidx_decode(qidx)
        sdims = [T.bonds[b].Qdegs[qin[b]] for b in range(nb)]
        qot = [qin[rsp[b]] for b in range(nb)]
        qot_off = sum(qot[b] * qot_acc[b] for b in range(nb))
        sbot = [1] * nb
        for b in range(nb - 1, 0, -1):
            sbot[rsp[b - 1]] = sbot[rsp[b]] * T.bonds[rsp[b]].Qdegs[qot[b]]
        ein, bin_c, sr, sc = L.subblock(qidx)
        eot, bot_c, osr, osc = Lo.subblock(qot_off)
        s = 1.0
        if swaps:
            bit = 0
            for b1, b2 in swaps:
                bit ^= T.bonds[b1].Qnums[qin[b1]].prtF & T.bonds[b2].Qnums[qin[b2]].prtF
            s = -1.0 if bit else 1.0
        meta.append([ein, bin_c, sc, eot, bot_c, osc])
        subd.append(sdims)
        acc.append(sbot)
        sign.append(s)
    return (np.array(meta, dtype=np.int64).reshape(-1, 6),
            np.array(subd, dtype=np.int64).reshape(-1, nb),
            np.array(acc, dtype=np.int64).reshape(-1, nb),
            np.array(sign, dtype=np.float64))


def run_moves(src: np.ndarray, dst: np.ndarray, rank: int, moves):
    meta, subd, acc, sign = moves
    src = np.ascontiguousarray(src)
    lib().ora_permute_moves(_ptr(src), _ptr(dst), rank, meta.shape[0],
                            _ptr(meta), _ptr(subd), _ptr(acc), _ptr(sign))


# ---------------------------------------------------------------------------
# contract  (S/tensor-network/lib/UniTensorTools.cpp:124-244)
# ---------------------------------------------------------------------------
def contract(Ta: UniTensor, Tb: UniTensor, fast: bool = False) -> UniTensor:
    """Permute A to (free | shared), B to (shared | free), then C_q = A_q B_q per common charge.

    Inputs are permuted in place and restored only when fast is False
    (UniTensorTools.cpp:178-179, 227-230).
    """
    if not (Ta.have_elem and Tb.have_elem):
        raise ValueError("Cannot perform contraction of two tensors before setting their elements.")
    if Ta is Tb:
        return contract(Ta, Tb.copy(), fast)
    if Ta.bonds and Tb.bonds:
        na, nbn = len(Ta.bonds), len(Tb.bonds)
        oldA, oldB = list(Ta.labels), list(Tb.labels)
        oldRA, oldRB = Ta.RBondNum, Tb.RBondNum
        newA, inter, newB, newC = [], [], [], []
        markB = [0] * nbn
        for a in range(na):
            match = False
            for b in range(nbn):
                if Ta.labels[a] == Tb.labels[b]:
                    markB[b] = 1
                    inter.append(Ta.labels[a])
                    newB.append(Tb.labels[b])
                    if Ta.bonds[a].dim != Tb.bonds[b].dim:
                        raise ValueError("Cannot contract two bonds having different dimensions")
                    match = True
                    break
            if not match:
                newA.append(Ta.labels[a])
                newC.append(Ta.labels[a])
        newA += inter
        for b in range(nbn):
            if markB[b] == 0:
                newB.append(Tb.labels[b])
                newC.append(Tb.labels[b])
        con = len(inter)
        Ta.permute(newA, na - con)
        Tb.permute(newB, con)
        cbonds = [Ta.bonds[i].copy() for i in range(na - con)] + [Tb.bonds[i].copy() for i in range(con, nbn)]
        Tc = UniTensor(cbonds)
        if cbonds:
            Tc.set_label(newC)
        for q, (ra, ca, offa) in Ta.layout.blocks.items():
            if q in Tb.layout.blocks:
                rb, cb_, offb = Tb.layout.blocks[q]
                rc, cc, offc = Tc.layout.blocks[q]
                if not (ra == rc and cb_ == cc and ca == rb):
                    raise ValueError("The dimensions the bonds to be contracted out are different.")
                A = Ta.elem[offa:offa + ra * ca].reshape(ra, ca)
                B = Tb.elem[offb:offb + rb * cb_].reshape(rb, cb_)
                Tc.elem[offc:offc + rc * cc] = (A @ B).reshape(-1)
        Tc.have_elem = True
        if con == 0:
            lab = oldA[:oldRA] + oldB[:oldRB] + oldA[oldRA:] + oldB[oldRB:]
            Tc.permute(lab, oldRA + oldRB)
        if not fast:
            Ta.permute(oldA, oldRA)
            Tb.permute(oldB, oldRB)
        return Tc
    if Ta.bonds:
        return Ta.copy().scale(Tb.elem[0])
    if Tb.bonds:
        return Tb.copy().scale(Ta.elem[0])
    return UniTensor(scalar=Ta.elem[0] * Tb.elem[0])


# ---------------------------------------------------------------------------
# Dense linear algebra with the reference's conventions
# ---------------------------------------------------------------------------
def otimes(Ta: UniTensor, Tb: UniTensor) -> UniTensor:
    """otimes (UniTensorTools.cpp:246-273): relabel so that nothing is shared
    (A's rows, B's rows, A's columns, B's columns) and contract."""
    T1, T2 = Ta.copy(), Tb.copy()
    l1 = [i if i < T1.RBondNum else T2.RBondNum + i for i in range(len(T1.bonds))]
    l2 = [i + T1.RBondNum if i < T2.RBondNum else i + len(T1.bonds) for i in range(len(T2.bonds))]
    T1.set_label(l1)
    T2.set_label(l2)
    return contract(T1, T2, True)


def rdotr(a: np.ndarray, a_shape, a_diag: bool, b: np.ndarray, b_shape, b_diag: bool):
    """RDotR (BlockTools.cpp:38-76): dense x dense -> matrixMul; diagonal x
    dense -> copy of b resized to (a rows, b cols), row i times a_i
    (diagRowMul); dense x diagonal -> copy of a resized, column j times b_j
    (diagColMul); diagonal x diagonal -> diagonal, entrywise product of the
    first min(|a|, |b|) entries.  A diagonal operand is its entry vector."""
    (ar, ac), (br, bc) = a_shape, b_shape
    if ac != br:
        raise ValueError("The dimensions of the two matrices do not match for matrix multiplication.")

    def resized(m, r, c):                       # Matrix::resize of a dense matrix (MatrixReal.cpp:289-359)
        out = np.zeros((r, c))
        rr, cc = min(r, m.shape[0]), min(c, m.shape[1])
        out[:rr, :cc] = m[:rr, :cc]
        return out
    if not a_diag and not b_diag:
        return np.asarray(a).reshape(ar, ac) @ np.asarray(b).reshape(br, bc), (ar, bc), False
    if a_diag and not b_diag:
        c = resized(np.asarray(b).reshape(br, bc), ar, bc)
        return c * np.asarray(a)[:ar, None], (ar, bc), False
    if not a_diag and b_diag:
        c = resized(np.asarray(a).reshape(ar, ac), ar, bc)
        return c * np.asarray(b)[None, :bc], (ar, bc), False
    n = min(ar, bc)
    out = np.zeros(n)
    k = min(len(a), len(b), n)
    out[:k] = np.asarray(a)[:k] * np.asarray(b)[:k]
    return out, (ar, bc), True


def svd(M: np.ndarray):
    """Block::svd -> matrixSVD (uni10_lapack_cpu.cpp:289-313).

    The reference hands the row-major MxN matrix to column-major dgesvd as
    its NxM transpose A^T = U' S V'^T, and reads U = V' (M x k) and
    VT = U'^T (k x N), k = min(M, N).  Same LAPACK routine (dgesvd).
    """
    import scipy.linalg as sla
    M = np.asarray(M, dtype=np.float64)
    Up, s, Vpt = sla.svd(M.T, full_matrices=False, lapack_driver="gesvd")
    return np.ascontiguousarray(Vpt.T), s, np.ascontiguousarray(Up.T)


def eigh(H: np.ndarray):
    """Block::eigh -> eigSyDecompose (uni10_lapack_cpu.cpp:131-152), dsyev('V','U').

    Eigenvalues ascending; eigenvectors returned as ROWS of a row-major matrix.
    """
    import scipy.linalg as sla
    w, v = sla.eigh(np.asarray(H, dtype=np.float64), driver="ev")
    return w, np.ascontiguousarray(v.T)


def exph(a: float, H: np.ndarray) -> np.ndarray:
    """exph (MatrixTools.cpp:109-121): U^T diag(exp(a*lambda)) U with U = eigenvector rows."""
    w, U = eigh(H)
    return U.T @ (np.exp(a * w)[:, None] * U)


def vector_norm(x: np.ndarray) -> float:
    return float(math.sqrt(float(np.dot(x, x))))


def lanczos_ev(A: np.ndarray, psi: np.ndarray, max_iter: int = 1000, err_tol: float = 5e-15):
    """lanczosEV (uni10_lapack_cpu.cpp:416-509) behind lanczosEigh (Block.cpp:474-506).

    Plain Lanczos from psi / ||psi|| without reorthogonalisation: dgemv (the
    row-major matrix times the vector, :444), ddot / daxpy (:445-447), norm
    and scale (:449-455); after every step the smallest eigenvalue of the
    tridiagonal matrix (dstev, :464-477; here SciPy's eigh_tridiagonal);
    stop on a relative change <= err_tol against max(|E0|, 1), on
    beta <= 1e-15, or at max_iter, never before 2 steps (:418-419, :431).
    Complex input follows the complex form (:905-1000): alpha = Re(v^H w).
    Returns (E0, psi, steps, converged).  The eigenvector's sign is left to
    dstev in the reference; here it is fixed so that <start|psi> >= 0.
    A space that closes after one step returns E0 = 0 and psi unchanged
    (:501-504).
    """
    import scipy.linalg as sla
    if not max_iter > 2:
        raise ValueError("Maximum iteration number should be set greater than 2.")
    dt = np.complex128 if (np.iscomplexobj(A) or np.iscomplexobj(psi)) else np.float64
    A = np.asarray(A, dtype=dt)
    N = A.shape[0]
    psi = np.asarray(psi, dtype=dt).reshape(-1)
    V = np.zeros((max_iter + 1, N), dtype=dt)
    nrm = lambda x: float(math.sqrt(float(np.vdot(x, x).real)))
    V[0] = psi * (1.0 / nrm(psi))
    As, Bs = [], []
    beta, e_diff, e0_old, it, conv = 1.0, 1.0, 0.0, 0, False
    beta_err = 1e-15

    def lowest(m, vec=False):
        d = np.array(As[:m])
        e = np.array(Bs[:m - 1])
        if vec:
            w, z = sla.eigh_tridiagonal(d, e, select="i", select_range=(0, 0))
            z = z[:, 0]
            return w[0], (z if z[0] >= 0 else -z)
        return sla.eigh_tridiagonal(d, e, eigvals_only=True, select="i", select_range=(0, 0))[0]

    while (((e_diff > err_tol) and it < max_iter) or it < 2) and beta > beta_err:
        w = A @ V[it] - (beta * V[it - 1] if it > 0 else 0.0)
        alpha = float(np.vdot(V[it], w).real)      # zdotc's real part for complex (:937-938)
        w = w - alpha * V[it]
        beta = nrm(w)
        As.append(alpha)
        if beta > beta_err:
            V[it + 1] = w * (1.0 / beta)
            Bs.append(beta if it < max_iter - 1 else 0.0)
        else:
            V[it + 1] = w
            Bs.append(0.0)
            conv = True
        it += 1
        if it > 1:
            d0 = lowest(it)
            base = abs(d0) if abs(d0) > 1 else 1.0
            e_diff = abs(d0 - e0_old) / base
            e0_old = d0
            if e_diff <= err_tol:
                conv = True
    if it > 1:
        lam, z = lowest(it, vec=True)
        return float(lam), z @ V[:it], it, conv
    return 0.0, psi.copy(), 1, conv


def heisenberg_ring(L: int, J: float = 1.0) -> np.ndarray:
    """H = J sum_i S_i . S_{i+1} on a periodic ring of L spins 1/2 (2^L x 2^L).
    Its ground energy at L = 4 is exactly -2 J."""
    h2 = heisenberg_two_site(1.0, J)
    n = 2 ** L
    H = np.zeros((n, n))
    for i in range(L):
        j = (i + 1) % L
        for s in range(n):
            bi, bj = (s >> (L - 1 - i)) & 1, (s >> (L - 1 - j)) & 1
            for a in range(2):
                for b in range(2):
                    v = h2[2 * a + b, 2 * bi + bj]
                    if v == 0.0:
                        continue
                    t = s
                    t = (t & ~(1 << (L - 1 - i))) | (a << (L - 1 - i))
                    t = (t & ~(1 << (L - 1 - j))) | (b << (L - 1 - j))
                    H[t, s] += v
    return H


# ---------------------------------------------------------------------------
# Network  (S/tensor-network/lib/Network.cpp:37-740)
# ---------------------------------------------------------------------------
def cal_elem_num(bonds: Sequence[Bond]) -> int:
    """Node::cal_elemNum (Network.cpp:137-219): element count of the block plan."""
    try:
        return Layout(bonds).elemNum
    except ValueError:
        return 0


class Node:
    def __init__(self, T: Optional[UniTensor] = None, bonds=None, labels=None):
        self.T = T
        if T is not None:
            self.bonds = [b.copy() for b in T.bonds]
            self.labels = list(T.labels)
            self.elemNum = T.elem_num()
            self.name = T.name
        else:
            self.bonds = bonds
            self.labels = labels
            self.elemNum = cal_elem_num(bonds)
            self.name = ""
        self.parent = self.left = self.right = None
        self.point = 0.0

    def _merge_bonds(self, nd):
        na, nbn = len(self.bonds), len(nd.bonds)
        markB = [0] * nbn
        cbonds, newl = [], []
        con = 0
        for a in range(na):
            match = False
            for b in range(nbn):
                if self.labels[a] == nd.labels[b]:
                    markB[b] = 1
                    match = True
                    con += 1
                    break
            if not match:
                newl.append(self.labels[a])
                cbonds.append(self.bonds[a].copy())
        for b in range(nbn):
            if markB[b] == 0:
                newl.append(nd.labels[b])
                cbonds.append(nd.bonds[b].copy())
        r = na - con
        for a in range(len(cbonds)):
            cbonds[a].change(BD_IN if a < r else BD_OUT)
        return cbonds, newl, con

    def contract(self, nd) -> "Node":
        """Node::contract (Network.cpp:63-100)."""
        cb, nl, _ = self._merge_bonds(nd)
        return Node(bonds=cb, labels=nl)

    def metric(self, nd) -> float:
        """Node::metric (Network.cpp:103-135): (elemA+elemB)/elemC, -1 if nothing is shared."""
        cb, _, con = self._merge_bonds(nd)
        if con == 0:
            return -1.0
        new = cal_elem_num(cb)
        return float(np.float32(np.float32(self.elemNum + nd.elemNum) / np.float32(new)))


class Network:
    """`.net` file parser, contraction-tree planner and post-order executor."""

    def __init__(self, fname: str):
        self.names: List[str] = []
        self.name2pos: Dict[str, int] = {}
        self.label_arr: List[List[int]] = []
        self.Rnums: List[int] = []
        self.brakets: List[int] = []
        self.order: List[int] = []
        self._fromfile(fname)
        n = len(self.label_arr) - 1
        self.tensors: List[Optional[UniTensor]] = [None] * n
        self.leafs: List[Optional[Node]] = [None] * n
        self.root: Optional[Node] = None
        self.load = False
        self.swapflags = [False] * n
        self.swaps_arr: List[list] = [[] for _ in range(n)]
        self.con_order: List[int] = []

    def _fromfile(self, fname):
        """Network::fromfile (Network.cpp:271-426)."""
        with open(fname) as f:
            # getline() followed by an eof() check drops a last line that has
            # no terminating newline (Network.cpp:288-291).
            lines = f.read().split("\n")[:-1]
        tar = set("1234567890-")
        ord_names: List[str] = []
        for raw in lines[:1000]:
            pos = raw.find(":")
            if pos < 0:
                break
            name = raw[:pos].strip()
            rest = raw[pos + 1:]
            if name == "ORDER":
                if "(" not in rest:
                    ord_names = [t for t in _split(rest, " ,;") if t]
                else:
                    ord_names, self.brakets = _parse_brackets(rest)
                    if sum(self.brakets) != 0:
                        raise ValueError("There are imbalance brackets when specifying the contraction order.")
                break
            self.name2pos[name] = len(self.names)
            self.names.append(name)
            labels, rnum = [], -1
            i = 0
            semi_seen = False
            while i < len(rest):
                ch = rest[i]
                if ch == ";":
                    semi_seen = True
                if ch in tar:
                    j = i
                    while j < len(rest) and rest[j] in tar:
                        j += 1
                    if rnum == -1 and semi_seen:
                        rnum = len(labels)
                    labels.append(_strtol(rest[i:j]))
                    i = j
                    continue
                i += 1
            self.label_arr.append(labels)
            self.Rnums.append(len(labels) if rnum == -1 else rnum)
        numT = len(self.names) - 1
        if numT < 0 or self.names[numT] != "TOUT":
            raise ValueError("Missing TOUT tensor.")
        if len(self.names) <= 2:
            raise ValueError("There must be at least two tensors in a tensor network.")
        if ord_names:
            if len(ord_names) != numT:
                raise ValueError("Some tensors are missing in the contraction order.")
            found = [False] * numT
            self.order = []
            for nm in ord_names:
                if nm not in self.name2pos:
                    raise ValueError("'%s' in the contraction order is not in the list of tensors above." % nm)
                p = self.name2pos[nm]
                if found[p]:
                    raise ValueError("'%s' appears more than once in the contraction order." % nm)
                found[p] = True
                self.order.append(p)
        else:
            self.order = list(range(numT))

    def put_tensor(self, key, T: UniTensor, force: bool = True):
        """Network::putTensor (Network.cpp:508-541): deep copy, relabelled."""
        idx = self.name2pos[key] if isinstance(key, str) else key
        if idx >= len(self.label_arr) - 1:
            raise IndexError("Index exceeds the number of the tensors in the list of network file.")
        if (not force) and self.load:
            self.destruct()
        if T.RBondNum != self.Rnums[idx]:
            raise ValueError("The number of in-coming bonds does not match with the tensor '%s'" % self.names[idx])
        t = T.copy()
        t.set_label(self.label_arr[idx])
        t.name = self.names[idx]
        self.tensors[idx] = t
        if self.leafs[idx] is None:
            self.leafs[idx] = Node(t)
        else:
            self.leafs[idx].T = t
            self.swapflags[idx] = False

    def put_tensor_T(self, key, T: UniTensor, force: bool = True):
        """Network::putTensorT (Network.cpp:576-591)."""
        t = T.copy().transpose()
        self.put_tensor(key, t, force)

    def destruct(self):
        """Network::destruct (Network.cpp:679-695): applied gates are applied
        again (undone), the gate lists cleared."""
        self.root = None
        for lf in self.leafs:
            if lf is not None:
                lf.parent = lf.left = lf.right = None
                lf.point = 0.0
        self.con_order = []
        for t in range(len(self.tensors)):
            if Qnum.fermionic and self.swapflags[t]:
                self.tensors[t].add_gate(self.swaps_arr[t])
                self.swapflags[t] = False
            self.swaps_arr[t] = []
        self.load = False

    def _find_con_ord(self, nd):
        """Network::findConOrd (Network.cpp:923-942): leaves in pre-order."""
        if nd is None or len(self.con_order) == len(self.tensors):
            return
        if nd.T is not None:
            pos = [i for i, t in enumerate(self.tensors) if t is nd.T]
            if not pos:
                raise RuntimeError("Fatal error(code = N2).")
            self.con_order.append(pos[0])
        self._find_con_ord(nd.left)
        self._find_con_ord(nd.right)

    def add_swap(self):
        """Network::addSwap (Network.cpp:944-993)."""
        n = len(self.leafs)
        self._find_con_ord(self.root)
        if len(self.con_order) != n:
            raise RuntimeError("Fatal error(code = N3).")
        for b1, b2 in rec_swap(self.con_order, list(range(n))):
            self.swaps_arr[b1] += self.tensors[b1].ex_swap(self.tensors[b2])
        for t in range(n):
            nb = len(self.tensors[t].bonds)
            recs: Dict[int, bool] = {}
            for a, b in self.swaps_arr[t]:
                h = min(a, b) * nb + max(a, b)
                recs[h] = not recs[h] if h in recs else True
            self.swaps_arr[t] = [(h // nb, h % nb) for h in sorted(recs) if recs[h]]

    def construct(self):
        """Network::construct (Network.cpp:428-506)."""
        if self.brakets:
            stack: List[Node] = []
            cnt = 0
            for b in self.brakets:
                if b < 0:
                    for _ in range(-b):
                        par = stack[-2].contract(stack[-1])
                        par.left, par.right = stack[-2], stack[-1]
                        par.left.parent = par.right.parent = par
                        stack = stack[:-2] + [par]
                        if len(stack) < 2:
                            break
                elif b == 0:
                    lf = self.leafs[self.order[cnt]]
                    if lf is None:
                        raise ValueError("Tensor '%s' has not yet been given." % self.names[self.order[cnt]])
                    stack.append(lf)
                    cnt += 1
            while len(stack) > 1:
                par = stack[-2].contract(stack[-1])
                par.left, par.right = stack[-2], stack[-1]
                par.left.parent = par.right.parent = par
                stack = stack[:-2] + [par]
            self.root = stack[0]
        else:
            for i in self.order:
                if self.leafs[i] is None:
                    raise ValueError("Tensor '%s' has not yet been given." % self.names[i])
                self._matching(self.leafs[i], self.root)
        tout = self.label_arr[-1]
        if len(self.root.labels) != len(tout) or any(l not in tout for l in self.root.labels):
            raise ValueError("The labels of the resulting tensor do not match with the labels of 'TOUT'")
        self.add_swap()
        self.load = True

    def _branch(self, sbj: Node, tar: Node):
        """Network::branch (Network.cpp:602-640)."""
        par = tar.contract(sbj)
        if sbj.parent is None:
            if tar.parent is not None:
                if tar.parent.left is tar:
                    tar.parent.left = par
                else:
                    tar.parent.right = par
                par.parent = tar.parent
            else:
                par.parent = None
                self.root = par
        else:
            gp = tar.parent.parent
            if gp is not None:
                if gp.left is tar.parent:
                    gp.left = par
                else:
                    gp.right = par
                par.parent = gp
            else:
                par.parent = None
                self.root = par
        par.left, par.right = tar, sbj
        tar.parent = par
        sbj.parent = par
        par.point = tar.metric(sbj)
        if sbj.parent.parent is not None:
            s = sbj.parent
            self._branch(s.parent.right, s.parent.left)

    def _matching(self, sbj: Node, tar: Optional[Node]):
        """Network::matching (Network.cpp:642-670)."""
        if tar is None:
            self.root = sbj
        elif tar.T is None:
            if sbj.metric(tar) > 0:
                tar_p = tar.point
                lft_p = sbj.metric(tar.left)
                rht_p = 0.0
                go = lft_p > tar_p
                if not go:
                    rht_p = sbj.metric(tar.right)
                    go = rht_p > tar_p
                if go:
                    if lft_p > rht_p:
                        self._matching(sbj, tar.left)
                    else:
                        self._matching(sbj, tar.right)
                else:
                    self._branch(sbj, tar)
            else:
                self._branch(sbj, tar)
        else:
            self._branch(sbj, tar)

    def _merge(self, nd: Node) -> UniTensor:
        """Network::merge (Network.cpp:719-739): post-order contract(fast=true)."""
        lt = self._merge(nd.left) if nd.left.T is None else nd.left.T
        rt = self._merge(nd.right) if nd.right.T is None else nd.right.T
        return contract(lt, rt, True)

    def launch(self, name: str = "") -> UniTensor:
        """Network::launch (Network.cpp:697-717)."""
        if not self.load:
            self.construct()
        for t in range(len(self.tensors)):
            if Qnum.fermionic and not self.swapflags[t]:
                self.tensors[t].add_gate(self.swaps_arr[t])
                self.swapflags[t] = True
        T = self._merge(self.root)
        tout = self.label_arr[-1]
        if len(tout) > 1:
            T.permute(tout, self.Rnums[-1])
        T.name = name
        return T

    def print_tree(self) -> str:
        """operator<< (Network.cpp:856-907): the tensor list, then the tree
        pre-order as "name(elemNum): labels," with "|   " per level."""
        out = ["\n"]
        for nm, labs, r in zip(self.names, self.label_arr, self.Rnums):
            line = nm + ": "
            if r:
                line += "i[" + ", ".join(map(str, labs[:r])) + "] "
            if len(labs) - r:
                line += "o[" + ", ".join(map(str, labs[r:])) + "]"
            out.append(line + "\n")
        out.append("\n")
        if all(lf is not None for lf in self.leafs):
            if not self.load:
                self.construct()

            def rec(n, layer):
                if n is None:
                    return
                head = "|   " * layer + ("%s(%d): " % (n.T.name, n.elemNum) if n.T is not None
                                         else "*(%d): " % n.elemNum)
                out.append(head + "".join("%d, " % l for l in n.labels) + "\n")
                rec(n.left, layer + 1)
                rec(n.right, layer + 1)
            rec(self.root, 0)
        else:
            out.append("\nSome tensors have not yet been given!\n\n")
        return "".join(out)

    def tree(self) -> list:
        """Nested (left, right) tuples of leaf names; for checking the planner."""
        def rec(n):
            if n.T is not None:
                return n.T.name
            return (rec(n.left), rec(n.right))
        if not self.load:
            self.construct()
        return rec(self.root)


def _split(s, dels):
    out, cur = [], ""
    for ch in s:
        if ch in dels:
            if cur:
                out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur:
        out.append(cur)
    return out


def _parse_brackets(s):
    """ORDER line with brackets (Network.cpp:318-352): names, and bracket
    markers (+n opens, -n closes, 0 = a tensor) in reading order."""
    names, br = [], []
    dels = " ,;*()"
    i = 0
    pending = ""
    while i < len(s):
        ch = s[i]
        if ch in dels:
            pending += ch
            i += 1
            continue
        j = i
        while j < len(s) and s[j] not in dels:
            j += 1
        closes = pending.count(")")
        opens = pending.count("(")
        if closes:
            br.append(-closes)
        if opens:
            br.append(opens)
        names.append(s[i:j])
        br.append(0)
        pending = ""
        i = j
    closes = pending.count(")")
    if closes:
        br.append(-closes)
    return names, br


def _strtol(tok: str) -> int:
    """strtol(…, 10) on a run of digits/minus signs: leading sign then digits."""
    i, sign = 0, 1
    if i < len(tok) and tok[i] in "+-":
        sign = -1 if tok[i] == "-" else 1
        i += 1
    j = i
    while j < len(tok) and tok[j].isdigit():
        j += 1
    return sign * int(tok[i:j]) if j > i else 0


# ---------------------------------------------------------------------------
# Old/new binary tensor format  (S/tensor-network/lib/UniTensor.cpp:486-582, 922-986)
# ---------------------------------------------------------------------------
def load_tensor(path: str) -> UniTensor:
    """Read a tensor dump; accepts the current format (r_flag, c_flag, name
    header) and the older one without it (the Swig example dumps)."""
    import struct
    b = open(path, "rb").read()

    def parse(o, header):
        def rd(fmt):
            nonlocal o
            v = struct.unpack_from(fmt, b, o)
            o += struct.calcsize(fmt)
            return v
        name = ""
        if header:
            rd("i"); rd("i")
            (nl,) = rd("i")
            name = b[o:o + nl].decode(errors="replace")
            o += nl
        (st,) = rd("i")
        (nb,) = rd("i")
        (qs,) = rd("Q")
        if qs != 12:
            raise ValueError("bad Qnum size")
        bonds = []
        for _ in range(nb):
            tp, nq = rd("ii")
            qn = [Qnum(*rd("iii")) for _ in range(nq)]
            dg = list(rd("%di" % nq))
            bonds.append(Bond.from_runs(tp, qn, dg))
        (nlab,) = rd("i")
        labels = list(rd("%di" % nlab))
        T = UniTensor(bonds, name)
        T.labels = labels
        if st & 2:
            (ne,) = rd("Q")
            T.set_elem(np.frombuffer(b, dtype=np.float64, count=ne, offset=o))
            o += 8 * ne
        if o != len(b):
            raise ValueError("trailing bytes")
        return T

    try:
        return parse(0, False)
    except Exception:
        return parse(0, True)


# ---------------------------------------------------------------------------
# Seeded synthetic inputs: SplitMix64 -> uniform[-1, 1)
# ---------------------------------------------------------------------------
_M64 = (1 << 64) - 1


def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    """x_{i} = SplitMix64 stream from `seed`; value = 2*((x >> 11) * 2^-53) - 1.

    Vectorised with uint64 wrap-around arithmetic; identical to the C++
    generator in the product (trim_amd/csrc/tn_rng.h).
    """
    i = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & _M64) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return 2.0 * u - 1.0


# ---------------------------------------------------------------------------
# iTEBD  (uni10/examples/iTEBD.cpp:148-216 with the .net files beside it)
# ---------------------------------------------------------------------------
def heisenberg_two_site(delta: float = 1.0, J: float = 1.0) -> np.ndarray:
    """Two-site spin-1/2 XXZ: J (Sx Sx + Sy Sy + delta Sz Sz), basis |uu>,|ud>,|du>,|dd>."""
    H = np.zeros((4, 4))
    H[0, 0] = H[3, 3] = 0.25 * delta * J
    H[1, 1] = H[2, 2] = -0.25 * delta * J
    H[1, 2] = H[2, 1] = 0.5 * J
    return H


def itebd_init(chi: int, d: int, seed: int):
    """Seeded starting point: Gamma_A, Gamma_B ~ U[-1,1) of shape (chi, d, chi)
    (seeds seed, seed+1); lambda_A, lambda_B = sorted-descending |U[-1,1)|
    (seeds seed+2, seed+3) normalised to unit 2-norm.  The reference's update
    works on ALa = Gamma_A lambda_A laid out (chi_left, chi_right | d), which is
    what this returns, with La/Lb as 1-D arrays."""
    ga = splitmix64_uniform(seed, chi * d * chi).reshape(chi, d, chi)
    gb = splitmix64_uniform(seed + 1, chi * d * chi).reshape(chi, d, chi)
    la = np.sort(np.abs(splitmix64_uniform(seed + 2, chi)))[::-1].copy()
    lb = np.sort(np.abs(splitmix64_uniform(seed + 3, chi)))[::-1].copy()
    la /= np.linalg.norm(la)
    lb /= np.linalg.norm(lb)
    ala = np.ascontiguousarray((ga * la[None, None, :]).transpose(0, 2, 1))
    blb = np.ascontiguousarray((gb * lb[None, None, :]).transpose(0, 2, 1))
    return ala, blb, la, lb


class ITEBD:
    """Plain (no-symmetry) iTEBD of the reference example, run through the
    oracle's UniTensor/Network restatement."""

    def __init__(self, chi: int, d: int, H: np.ndarray, tau: float, seed: int, netdir: str):
        self.chi, self.d, self.tau = chi, d, tau
        ala, blb, la, lb = itebd_init(chi, d, seed)
        self.H = H
        bH = [Bond(BD_IN, d), Bond(BD_IN, d), Bond(BD_OUT, d), Bond(BD_OUT, d)]
        self.Ht = UniTensor(bH, "H")
        self.Ht.put_block(H)
        self.U = UniTensor(bH, "U")
        self.U.put_block(exph(-tau, H))
        b3 = [Bond(BD_IN, chi), Bond(BD_IN, chi), Bond(BD_OUT, d)]
        self.ALa = UniTensor(b3, "ALa", elem=ala.reshape(-1))
        self.BLb = UniTensor(b3, "BLb", elem=blb.reshape(-1))
        self.La, self.Lb = la, lb
        self.net = Network(os.path.join(netdir, "itebd.net"))
        self.netA = Network(os.path.join(netdir, "updateA.net"))
        self.mps = Network(os.path.join(netdir, "MPS.net"))
        self.meas = Network(os.path.join(netdir, "measure.net"))
        self.trunc_err = 0.0

    def _update(self, ALa, BLb, La, Lb):
        """iTEBD.cpp:148-173 (one two-site update); returns (new ALa, new BLb, new La)."""
        self.net.put_tensor("ALa", ALa)
        self.net.put_tensor("BLb", BLb)
        self.net.put_tensor("expH", self.U)
        C = self.net.launch()
        Theta = UniTensor(C.bonds, "Theta")
        Theta.put_block(Lb[:, None] * C.get_block())
        Theta.permute_row(2)
        U, s, VT = svd(Theta.get_block())
        dim = min(self.chi, s.shape[0])
        self.trunc_err = float(np.sum(s[dim:] ** 2))
        S = s[:dim]
        nrm = float(np.sqrt(np.sum(S * S)))
        La_new = S / nrm
        b3 = [b.copy() for b in ALa.bonds]
        b3[0] = Bond(BD_IN, dim)
        BLb_new = UniTensor(b3)
        BLb_new.set_elem(VT.reshape(-1)[: BLb_new.elem_num()])
        self.netA.put_tensor("BLb", BLb_new)
        self.netA.put_tensor("C", C)
        ALa_new = self.netA.launch()
        ALa_new.scale(1.0 / nrm)
        return ALa_new, BLb_new, La_new

    def step(self):
        """One 'step' = update(A,B) then update(B,A), as iTEBD.cpp:128-129."""
        self.ALa, self.BLb, self.La = self._update(self.ALa, self.BLb, self.La, self.Lb)
        self.BLb, self.ALa, self.Lb = self._update(self.BLb, self.ALa, self.Lb, self.La)

    def _expectation(self, L, R):
        """iTEBD.cpp:205-216."""
        self.mps.put_tensor("L", L)
        self.mps.put_tensor("R", R)
        psi = self.mps.launch()
        n = np.linalg.norm(psi.get_block())
        self.meas.put_tensor("bra", psi)
        self.meas.put_tensor_T("ket", psi)
        self.meas.put_tensor("Op", self.Ht)
        return self.meas.launch()[0] / (n * n)

    def energy(self) -> float:
        """iTEBD.cpp:189-203: mean of the two bond energies."""
        A1 = self.ALa.copy()
        A1.permute_row(1)
        A1.put_block(self.Lb[:, None] * A1.get_block())
        A1.permute_row(2)
        B1 = self.BLb.copy()
        B1.permute_row(1)
        B1.put_block(self.La[:, None] * B1.get_block())
        B1.permute_row(2)
        v = self._expectation(A1, self.BLb)
        v += self._expectation(B1, self.ALa)
        return v / 2.0


# ---------------------------------------------------------------------------
# Spin-1 chains and U(1)-symmetric iTEBD (the batched workloads' symmetric
# chains).  The reference's iTEBD example is dense; its symmetric tensors
# (UniTensorReal.cpp:1169-1385) hold one block per total charge, and
# Block::svd (BlockReal.cpp:226-250) factorises one block.  A U(1) update is
# the dense update of iTEBD.cpp:148-173 with theta block-diagonal in the
# charge of the middle bond: every charge block is factorised on its own and
# keeps a fixed number of singular values (the bond's sector dimension), so
# every tensor keeps its block shapes from step to step.
# ---------------------------------------------------------------------------
def xxz_two_site_spin1(delta: float = 1.0, J: float = 1.0) -> np.ndarray:
    """Two-site spin-1 XXZ J (Sx Sx + Sy Sy + delta Sz Sz) = J ((S+S- + S-S+)/2
    + delta Sz Sz), basis |s t> with s, t in (+1, 0, -1), s the slower index."""
    sz = np.diag([1.0, 0.0, -1.0])
    sp = np.sqrt(2.0) * np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, 0.0]])
    sm = sp.T
    return J * (0.5 * (np.kron(sp, sm) + np.kron(sm, sp)) + delta * np.kron(sz, sz))


def xxz_two_site_d(delta: float, d: int) -> np.ndarray:
    if d == 2:
        return heisenberg_two_site(delta)
    if d == 3:
        return xxz_two_site_spin1(delta)
    raise ValueError("d must be 2 or 3")


def phys_charges(d: int) -> List[int]:
    """2 Sz of the physical states, in basis order: (+1, -1) or (+2, 0, -2)."""
    return {2: [1, -1], 3: [2, 0, -2]}[d]


def _largest_remainder(total: int, w: np.ndarray) -> np.ndarray:
    """Integers >= 1 summing to total, proportional to w (largest remainder;
    ties to the lower index)."""
    n = len(w)
    x = w / w.sum() * (total - n)
    base = np.floor(x).astype(np.int64)
    rem = x - base
    short = (total - n) - int(base.sum())
    order = sorted(range(n), key=lambda i: (-rem[i], i))
    for i in order[:short]:
        base[i] += 1
    return base + 1


def u1_bond_sectors(chi: int, d: int, qmax: int = 12):
    """Charges (ascending) and dimensions of the two bonds of a U(1) chain of
    bond dimension chi: bond 1 (lambda_B, left of site A) holds even charges,
    bond 0 (lambda_A) the charges q + s reachable through one site, both
    within [-qmax, qmax].  Dimensions follow a Gaussian of width
    sigma = 2 (1 + log2(chi) / 4) over the charge, rounded by largest
    remainder with at least 1 per sector, on the charges where the Gaussian
    share is at least 1/2; a sector never holds more states than the theta
    block that produces it (sum over s of the outer bond's sector q - s),
    excess being moved to the sector with the most room."""
    phys = phys_charges(d)
    sigma = 2.0 * (1.0 + math.log2(chi) / 4.0)

    def pick(cands):
        cands = np.array(sorted(cands))
        w = np.exp(-cands.astype(np.float64) ** 2 / (2 * sigma * sigma))
        keep = chi * w / w.sum() >= 0.5
        keep[np.argmax(w)] = True
        cands, w = cands[keep], w[keep]
        if len(cands) > chi:
            order = np.argsort(-w, kind="stable")[:chi]
            cands, w = cands[np.sort(order)], w[np.sort(order)]
        return [int(c) for c in cands], [int(v) for v in _largest_remainder(chi, w)]

    q1, n1 = pick([q for q in range(-qmax, qmax + 1) if q % 2 == 0])
    q0, n0 = pick({q + s for q in q1 for s in phys if abs(q + s) <= qmax})
    bonds = [[q0, n0], [q1, n1]]
    for _ in range(64):
        changed = False
        for p in (0, 1):
            qR, nR = bonds[p]
            qL, nL = bonds[1 - p]
            dimL = dict(zip(qL, nL))
            room = [sum(dimL.get(q - s, 0) for s in phys) for q in qR]
            for i in range(len(qR)):
                while nR[i] > room[i]:
                    nR[i] -= 1
                    j = max(range(len(qR)), key=lambda k: (room[k] - nR[k], -abs(qR[k]), -qR[k]))
                    if room[j] - nR[j] <= 0:
                        raise ValueError("u1_bond_sectors: chi=%d does not fit the charge range" % chi)
                    nR[j] += 1
                    changed = True
        if not changed:
            break
    return (bonds[0][0], bonds[0][1]), (bonds[1][0], bonds[1][1])


class ITEBDU1:
    """U(1)-symmetric iTEBD on a spin-1/2 (d = 2) or spin-1 (d = 3) XXZ chain,
    restated with dense arrays and charge masks (the GPU engine stores the
    blocks).  Tensor p (Gamma_p lambda_p, as ALa/BLb of iTEBD.cpp) has left
    bond 1 - p and right bond p, entries only where q_right = q_left + s.
    The seeded start: Gamma_p entries splitmix64(seed + p) at the dense index
    (a d + s) chi + m, masked; lambda_p = |splitmix64(seed + 2 + p)| sorted
    descending within each sector, normalised to unit 2-norm."""

    def __init__(self, chi: int, d: int, delta: float, tau: float, seed: int, netdir: str, qmax: int = 12):
        self.chi, self.d, self.tau, self.delta, self.netdir = chi, d, tau, delta, netdir
        self.phys = np.array(phys_charges(d))
        self.sect = list(u1_bond_sectors(chi, d, qmax))
        self.q = [np.repeat(np.array(s[0]), s[1]) for s in self.sect]   # charge of every bond index
        self.H = xxz_two_site_d(delta, d)
        self.G = exph(-tau, self.H).reshape(d, d, d, d)
        self.lam = []
        for p in (0, 1):
            raw = np.abs(splitmix64_uniform(seed + 2 + p, chi))
            lam = np.empty(chi)
            o = 0
            for n in self.sect[p][1]:
                lam[o:o + n] = np.sort(raw[o:o + n])[::-1]
                o += n
            self.lam.append(lam / np.linalg.norm(lam))
        self.X = []
        for p in (0, 1):
            qL, qR = self.q[1 - p], self.q[p]
            g = splitmix64_uniform(seed + p, chi * d * chi).reshape(chi, d, chi)
            mask = (qR[None, None, :] == qL[:, None, None] + self.phys[None, :, None])
            g = np.where(mask, g, 0.0)
            self.X.append(np.ascontiguousarray((g * self.lam[p][None, None, :]).transpose(0, 2, 1)))
        self.trunc_err = 0.0

    def _update(self, p: int, keep_all: bool = False):
        """iTEBD.cpp:148-173 with theta factorised per charge block of the
        middle bond p (Block::svd, BlockReal.cpp:226-250); the outer bond is
        1 - p.  keep_all keeps every singular value (the bond p sectors then
        take the block sizes)."""
        d, phys = self.d, self.phys
        X, Y = self.X[p], self.X[1 - p]
        qL = self.q[1 - p]
        cL = len(qL)
        T1 = np.einsum("ams,mbt->asbt", X, Y)
        C = np.einsum("asbt,stuv->aubv", T1, self.G)
        Th = self.lam[1 - p][:, None, None, None] * C
        Cm = C.reshape(cL * d, cL * d)
        Thm = Th.reshape(cL * d, cL * d)
        rowq = (qL[:, None] + phys[None, :]).reshape(-1)
        colq = (qL[:, None] - phys[None, :]).reshape(-1)
        charges, dims = self.sect[p]
        newS, newVT, newq, disc = [], [], [], 0.0
        for q, k in zip(charges, dims):
            rows = np.nonzero(rowq == q)[0]
            cols = np.nonzero(colq == q)[0]
            blk = Thm[np.ix_(rows, cols)]
            s, vt = svd(blk)[1:]
            kk = len(s) if keep_all else k
            if kk > len(s):
                raise ValueError("sector %d keeps %d of %d singular values" % (q, kk, len(s)))
            disc += float(np.sum(s[kk:] ** 2))
            full_vt = np.zeros((kk, cL * d))
            full_vt[:, cols] = vt[:kk]
            newS.append(s[:kk])
            newVT.append(full_vt)
            newq.append(np.full(kk, q))
        S = np.concatenate(newS)
        VT = np.concatenate(newVT)
        nrm = float(np.sqrt(np.sum(S * S)))
        if keep_all:
            self.sect[p] = [list(charges), [len(s) for s in newS]]
            self.q[p] = np.concatenate(newq)
        cR = len(S)
        self.lam[p] = S / nrm
        self.X[1 - p] = VT.reshape(cR, cL, d)
        self.X[p] = np.ascontiguousarray((Cm @ VT.T).reshape(cL, d, cR).transpose(0, 2, 1)) / nrm
        self.trunc_err = disc

    def step(self):
        """update(A,B) then update(B,A), as iTEBD.cpp:128-129."""
        self._update(0)
        self._update(1)

    def dense(self):
        """(ALa, BLb, La, Lb) in the dense layout of ITEBD."""
        return self.X[0], self.X[1], self.lam[0], self.lam[1]

    def energy(self) -> float:
        """Energy per site of the state, through ITEBD's networks
        (iTEBD.cpp:189-216) on the dense tensors."""
        E = ITEBD(1, self.d, self.H, self.tau, 1, self.netdir)
        c = self.chi
        b3 = [Bond(BD_IN, c), Bond(BD_IN, c), Bond(BD_OUT, self.d)]
        E.ALa = UniTensor(b3, "ALa", elem=self.X[0].reshape(-1))
        E.BLb = UniTensor(b3, "BLb", elem=self.X[1].reshape(-1))
        E.La, E.Lb = self.lam[0], self.lam[1]
        return E.energy()
