"""Pins the CPU oracle (oracle/uni10_oracle.py, oracle/uni10_cpu.c) to the
reference's golden data, its unit-test known answers and closed forms.

Nothing here re-types an oracle formula as its own check: every expected
value comes from a reference fixture/test, a textbook routine
(numpy.transpose, dense tensordot, LAPACK), or mathematics (closed forms,
invariants, brute force).
"""
import math
import os

import numpy as np
import pytest

from netcases import (CHAIN3_NET, FERMI_ORDERS, fermi_tensor, NETWORK_H_NET, NETWORK_H_PRINT, network_h_bonds,
                      network_h_fermionic_bonds)


# ---------------------------------------------------------------------------
# Qnum known answers (uni10/test/testQnum.cpp)
def test_qnum_known_answers(ora):
    q = ora.Qnum(100, ora.PRT_EVEN, ora.PRTF_ODD)
    assert q.hash() == 401                                   # testQnum.cpp:92-95
    q1 = -ora.Qnum(-1)
    assert (q1.U1, q1.prt) == (1, ora.PRT_EVEN)              # :97-108
    q2 = -ora.Qnum(3, ora.PRT_ODD, ora.PRTF_EVEN)
    assert (q2.U1, q2.prt, q2.prtF) == (-3, ora.PRT_ODD, ora.PRTF_EVEN)
    assert ora.Qnum(-1) < ora.Qnum(3)                        # :110-122
    assert ora.Qnum(0, 0, 0) < ora.Qnum(0, 0, 1)
    assert ora.Qnum(0, 0, 0) < ora.Qnum(0, 1, 0)
    a = ora.Qnum(3, ora.PRT_ODD, ora.PRTF_EVEN) * ora.Qnum(2, ora.PRT_ODD, ora.PRTF_ODD)
    assert (a.U1, a.prt, a.prtF) == (5, ora.PRT_EVEN, ora.PRTF_ODD)   # :148-160
    b = ora.Qnum(-1, ora.PRT_ODD) * ora.Qnum(2, ora.PRT_ODD)
    assert (b.U1, b.prt, b.prtF) == (1, ora.PRT_EVEN, ora.PRTF_EVEN)
    assert not (ora.Qnum(0, 0, 0) == ora.Qnum(0, 1, 0))
    with pytest.raises(ValueError):
        ora.Qnum(1000)
    with pytest.raises(ValueError):
        ora.Qnum(-1000)


# ---------------------------------------------------------------------------
# Bond known answers (uni10/test/testBond.cpp)
def test_bond_known_answers(ora):
    Q = ora.Qnum
    bd = ora.Bond(ora.BD_OUT, [Q(1), Q(1), Q(0), Q(0), Q(0), Q(-1)])
    assert bd.dim == 6
    assert [q.U1 for q in bd.Qlist()] == [1, 1, 0, 0, 0, -1]
    assert [(q.U1, d) for q, d in bd.degeneracy().items()] == [(-1, 1), (0, 3), (1, 2)]
    bd1 = ora.Bond(ora.BD_IN, [Q(1), Q(1), Q(0), Q(0), Q(0), Q(-1)])
    bd2 = ora.Bond(ora.BD_IN, [Q(1), Q(0), Q(0), Q(-1)])
    bd2.combine(bd1)
    assert bd2.type == ora.BD_IN and bd2.dim == 24            # testBond.cpp:101-160
    assert [(q.U1, d) for q, d in bd2.degeneracy().items()] == [(-2, 1), (-1, 5), (0, 9), (1, 7), (2, 2)]
    b = ora.Bond(ora.BD_IN, 100)
    b.change(ora.BD_OUT)
    assert b.type == ora.BD_OUT
    # non-adjacent repeats stay separate runs
    r = ora.Bond(ora.BD_IN, [Q(1), Q(0), Q(1)])
    assert [q.U1 for q in r.Qnums] == [1, 0, 1] and r.Qdegs == [1, 1, 1]
    with pytest.raises(ValueError):
        ora.Bond(ora.BD_IN, [])


def test_combine_bonds_brute_force(ora):
    """combine(tp, bonds): the fused charge of every state of the product space
    in row-major order equals the sum of the per-bond charges (brute force)."""
    Q = ora.Qnum
    bds = [ora.Bond(ora.BD_IN, [Q(1), Q(-1)]), ora.Bond(ora.BD_OUT, [Q(2), Q(0), Q(0)]),
           ora.Bond(ora.BD_IN, [Q(1), Q(0), Q(-1)])]
    c = ora.combine_bonds(ora.BD_IN, bds)
    expect = []
    for s0 in bds[0].Qlist():
        for s1 in bds[1].copy().change(ora.BD_IN).Qlist():
            for s2 in bds[2].Qlist():
                expect.append(s0.U1 + s1.U1 + s2.U1)
    assert [q.U1 for q in c.Qlist()] == expect


# ---------------------------------------------------------------------------
# Block layout: golden tensor dumps of the reference (uni10/Swig/examples)
HEIS_S1 = np.array([
    1, 0, 0, 0, 0, 0, 0, 0, 0,
    0, 0, 0, 1, 0, 0, 0, 0, 0,
    0, 0, -1, 0, 1, 0, 0, 0, 0,
    0, 1, 0, 0, 0, 0, 0, 0, 0,
    0, 0, 1, 0, 0, 0, 1, 0, 0,
    0, 0, 0, 0, 0, 0, 0, 1, 0,
    0, 0, 0, 0, 1, 0, -1, 0, 0,
    0, 0, 0, 0, 0, 1, 0, 0, 0,
    0, 0, 0, 0, 0, 0, 0, 0, 1], dtype=np.float64)   # egU1.cpp:31-41


def test_grouping_matches_egU1_dump(ora, golden):
    Q = ora.Qnum
    qs = [Q(1), Q(0), Q(-1)]
    bi, bo = ora.Bond(ora.BD_IN, qs), ora.Bond(ora.BD_OUT, qs)
    H = ora.UniTensor([bi, bi, bo, bo], "H_U1")
    H.set_raw_elem(HEIS_S1)
    G = ora.load_tensor(os.path.join(golden, "egU1_H_U1"))
    assert [b == g for b, g in zip(H.bonds, G.bonds)] == [True] * 4
    assert H.elem.tobytes() == G.elem.tobytes()
    assert [q.U1 for q in H.block_qnums()] == [-2, -1, 0, 1, 2]
    assert [H.layout.blocks[q][:2] for q in H.block_qnums()] == [(1, 1), (2, 2), (3, 3), (2, 2), (1, 1)]
    # getRawElem inverts setRawElem (all nonzeros of this H sit in blocks)
    assert np.array_equal(H.get_raw_elem().reshape(-1), HEIS_S1)
    # at() on every index
    raw = HEIS_S1.reshape(3, 3, 3, 3)
    for idx in np.ndindex(3, 3, 3, 3):
        assert H.at(idx) == raw[idx]


def test_transpose_matches_egU3_dump(ora, golden):
    W = ora.load_tensor(os.path.join(golden, "egU3_W"))
    WT = ora.load_tensor(os.path.join(golden, "egU3_WT"))
    assert W.elem_num() == 12 and [b.dim for b in W.bonds] == [6, 3, 3]
    T = W.copy().transpose()
    assert T.elem.tobytes() == WT.elem.tobytes()
    assert all(a == b for a, b in zip(T.bonds, WT.bonds))
    # W is an isometry block by block (orthoRand): rows orthonormal
    for q in W.block_qnums():
        blk = W.get_block(q)
        assert np.allclose(blk @ blk.T, np.eye(blk.shape[0]), atol=1e-14)


def test_block_layout_brute_force(ora):
    """Every element of setRawElem's buffer lands at the position of its
    (charge block, row, col) computed by brute force from the row-major
    odometers over the in- and out-bonds (independent statement of the layout)."""
    Q = ora.Qnum
    bonds = [ora.Bond(ora.BD_IN, [Q(1), Q(-1), Q(1)]), ora.Bond(ora.BD_IN, [Q(0), Q(2), Q(2)]),
             ora.Bond(ora.BD_OUT, [Q(1), Q(1), Q(3), Q(-1)])]
    T = ora.UniTensor(bonds)
    dims = [b.dim for b in bonds]
    raw = np.arange(1, np.prod(dims) + 1, dtype=np.float64).reshape(dims)
    T.set_raw_elem(raw)
    # brute force: group row states by charge in run-length-odometer order
    def states(bl):
        runs = []
        for b in bl:
            runs.append([(q, list(range(o, o + d))) for q, d, o in zip(b.Qnums, b.Qdegs, b.offsets)])
        out = []
        import itertools
        for combo in itertools.product(*runs):
            q = sum(c[0].U1 for c in combo)
            for idx in itertools.product(*[c[1] for c in combo]):
                out.append((q, idx))
        return out
    rows = states(bonds[:2])
    cols = states(bonds[2:])
    expect = []
    for q in sorted(set(r[0] for r in rows) & set(c[0] for c in cols)):
        for r in [x for x in rows if x[0] == q]:
            for c in [x for x in cols if x[0] == q]:
                expect.append(raw[r[1] + c[1]])
    assert T.elem.tolist() == expect


# ---------------------------------------------------------------------------
# Dense permute: the C odometer vs numpy.transpose vs the _reshapeElem formula
@pytest.mark.parametrize("seed", range(12))
def test_dense_permute_three_ways(ora, seed):
    rng = np.random.default_rng(seed)
    r = int(rng.integers(1, 7))
    dims = np.array(rng.integers(1, 6, size=r), dtype=np.int64)
    perm = np.array(rng.permutation(r), dtype=np.int32)
    n = int(np.prod(dims))
    x = rng.standard_normal(n)
    a, b = np.empty(n), np.empty(n)
    ora.lib().ora_permute_dense(ora._ptr(x), ora._ptr(a), r, ora._ptr(dims), ora._ptr(perm))
    ora.lib().ora_reshape_elem(ora._ptr(x), ora._ptr(b), r, ora._ptr(dims), ora._ptr(perm))
    c = np.transpose(x.reshape(dims), perm).reshape(-1)
    assert a.tobytes() == c.tobytes() and b.tobytes() == c.tobytes()


def test_dense_permute_catches_planted_bug(ora):
    """A transposed perm (the inverse permutation) must be caught by the check."""
    dims = np.array([2, 3, 4], np.int64)
    perm = np.array([1, 2, 0], np.int32)
    inv = np.argsort(perm).astype(np.int32)
    x = np.arange(24, dtype=np.float64)
    a = np.empty(24)
    ora.lib().ora_permute_dense(ora._ptr(x), ora._ptr(a), 3, ora._ptr(dims), ora._ptr(inv))
    assert a.tobytes() != np.transpose(x.reshape(2, 3, 4), perm).reshape(-1).tobytes()


# ---------------------------------------------------------------------------
# Symmetric permute: densify-and-compare
def _random_u1(ora, seed, n_in=2, n_out=2):
    rng = np.random.default_rng(seed)
    Q = ora.Qnum
    bonds = []
    for i in range(n_in + n_out):
        tp = ora.BD_IN if i < n_in else ora.BD_OUT
        charges = sorted(set(int(c) for c in rng.integers(-3, 4, size=int(rng.integers(1, 4)))))
        qs = []
        for c in charges:
            qs += [Q(c)] * int(rng.integers(1, 4))
        if rng.random() < 0.3:     # a non-adjacent repeat
            qs += [qs[0]]
        bonds.append(ora.Bond(tp, qs))
    try:
        T = ora.UniTensor(bonds)
    except ValueError:
        return None
    T.set_raw_elem(rng.uniform(-1, 1, size=[b.dim for b in bonds]))
    return T


@pytest.mark.parametrize("seed", range(20))
def test_symmetric_permute_densify(ora, seed):
    rng = np.random.default_rng(100 + seed)
    T = _random_u1(ora, seed)
    if T is None:
        pytest.skip("random bonds carry no common charge")
    nb = len(T.bonds)
    perm = [int(v) for v in rng.permutation(nb)]
    rb = int(rng.integers(0, nb + 1))
    before = T.get_raw_elem().reshape([b.dim for b in T.bonds])
    labels = [T.labels[p] for p in perm]
    try:
        P = T.copy().permute(labels, rb)
    except ValueError:
        pytest.skip("permuted bonds carry no common charge")
    after = P.get_raw_elem().reshape([b.dim for b in P.bonds])
    # elements that survive (inside blocks of both layouts) are moved exactly
    mask_in = np.zeros(before.shape, bool)
    T2 = T.copy()
    T2.set_elem(np.ones(T2.elem_num()))
    mask_in = T2.get_raw_elem().reshape(before.shape) != 0
    expect = np.transpose(np.where(mask_in, before, 0.0), perm)
    assert np.array_equal(after, expect)
    # and back
    back = P.copy().permute(T.labels, T.RBondNum)
    assert back.elem.tobytes() == T.elem.tobytes()


def _perm_sign_by_cycles(order):
    seen, sign = [False] * len(order), 1
    for s in range(len(order)):
        length, i = 0, s
        while not seen[i]:
            seen[i] = True
            i = order[i]
            length += 1
        if length and length % 2 == 0:
            sign = -sign
    return sign


def test_rec_swap_lists_exactly_the_inverted_pairs(ora):
    """Bubble sort swaps every inverted pair of the order exactly once, so the
    swap list that sets the fermionic signs is the inversion set (mapped
    through ordF); with every bond odd its length has the parity of the
    permutation, whose sign the cycle decomposition gives independently."""
    rng = np.random.default_rng(11)
    for n in range(1, 8):
        for _ in range(30):
            order = [int(v) for v in rng.permutation(n)]
            ordF = [int(v) for v in rng.permutation(n)]
            got = ora.rec_swap(order, ordF)
            want = {(ordF[order[j]], ordF[order[i]])
                    for i in range(n) for j in range(i + 1, n) if order[i] > order[j]}
            assert len(got) == len(want) and set(got) == want
            assert (-1) ** len(got) == _perm_sign_by_cycles(order)


def test_contract_vs_tensordot(ora):
    for seed in range(6):
        rng = np.random.default_rng(seed)
        Q = ora.Qnum
        chi = ora.Bond(ora.BD_IN, [Q(-1)] * 2 + [Q(1)] * 3)
        phys = ora.Bond(ora.BD_OUT, [Q(1), Q(-1)])
        A = ora.UniTensor([chi, phys, ora.Bond(ora.BD_OUT, [Q(0)] * 2 + [Q(2)] * 2 + [Q(-2)])])
        A.set_raw_elem(rng.uniform(-1, 1, [b.dim for b in A.bonds]))
        A.set_label([1, 2, 3])
        B = ora.UniTensor([ora.Bond(ora.BD_IN, [Q(0)] * 2 + [Q(2)] * 2 + [Q(-2)]),
                           ora.Bond(ora.BD_OUT, [Q(1), Q(-1)]), ora.Bond(ora.BD_OUT, [Q(-1)] * 3 + [Q(1)])])
        B.set_raw_elem(rng.uniform(-1, 1, [b.dim for b in B.bonds]))
        B.set_label([3, 4, 5])
        Ad = A.get_raw_elem().reshape([b.dim for b in A.bonds])
        Bd = B.get_raw_elem().reshape([b.dim for b in B.bonds])
        C = ora.contract(A, B)
        Cd = C.get_raw_elem().reshape([b.dim for b in C.bonds])
        ref = np.tensordot(Ad, Bd, axes=([2], [0]))
        assert np.allclose(Cd, ref, atol=1e-14, rtol=0)
        assert A.labels == [1, 2, 3] and B.labels == [3, 4, 5]


def test_contract_dense_c2_shape(ora):
    """The contraction-and-permute workload at small chi, against tensordot."""
    chi = 8
    A = ora.UniTensor([ora.Bond(ora.BD_IN, chi), ora.Bond(ora.BD_IN, 2), ora.Bond(ora.BD_OUT, 2),
                       ora.Bond(ora.BD_OUT, chi)])
    B = ora.UniTensor([ora.Bond(ora.BD_IN, chi), ora.Bond(ora.BD_IN, 2), ora.Bond(ora.BD_OUT, 2),
                       ora.Bond(ora.BD_OUT, chi)])
    a = ora.splitmix64_uniform(1, A.elem_num())
    b = ora.splitmix64_uniform(2, B.elem_num())
    A.set_elem(a)
    B.set_elem(b)
    A.set_label([0, 1, 2, 3])
    B.set_label([3, 2, 4, 5])
    C = ora.contract(A, B)
    assert C.labels == [0, 1, 4, 5]
    ref = np.tensordot(a.reshape(chi, 2, 2, chi), b.reshape(chi, 2, 2, chi), axes=([2, 3], [1, 0]))
    assert np.allclose(C.elem.reshape(ref.shape), ref, rtol=0, atol=1e-13)
    before = C.elem.reshape(chi, 2, 2, chi).copy()
    C.permute([5, 1, 0, 4], 2)
    # labels (0,1,4,5) -> (5,1,0,4): output axis j is input axis (3,1,0,2)[j]
    assert C.elem.reshape(chi, 2, chi, 2).tobytes() == np.transpose(before, (3, 1, 0, 2)).tobytes()


# ---------------------------------------------------------------------------
# SVD / eigh conventions
def test_svd_golden_dump(ora, golden):
    U = np.fromfile(os.path.join(golden, "mat_U"), dtype=np.float64).reshape(4, 4)
    S = np.fromfile(os.path.join(golden, "mat_Sigma"), dtype=np.float64)
    VT = np.fromfile(os.path.join(golden, "mat_VT"), dtype=np.float64).reshape(4, 5)
    assert np.all(np.diff(S) < 0)
    assert np.allclose(U.T @ U, np.eye(4), atol=1e-15 * 10)
    assert np.allclose(VT @ VT.T, np.eye(4), atol=1e-15 * 10)
    M = U @ np.diag(S) @ VT
    assert M.min() >= -1e-12 and M.max() < 1.0     # Matrix::randomize draws from [0,1)
    u, s, vt = ora.svd(M)
    assert u.shape == (4, 4) and s.shape == (4,) and vt.shape == (4, 5)
    assert np.allclose(s, S, rtol=1e-14, atol=0)
    for j in range(4):
        sg = np.sign(u[:, j] @ U[:, j])
        assert np.allclose(sg * u[:, j], U[:, j], atol=1e-13)
        assert np.allclose(sg * vt[j], VT[j], atol=1e-13)


def test_heisenberg_gate_closed_form(ora):
    H = ora.heisenberg_two_site(1.0)
    w, U = ora.eigh(H)
    assert np.allclose(w, [-0.75, 0.25, 0.25, 0.25], atol=1e-15)
    tau = 0.01
    G = ora.exph(-tau, H)
    # exp(-tau H) = e^{-tau/4} P_triplet + e^{3 tau/4} P_singlet
    s = np.array([0, 1, -1, 0]) / math.sqrt(2)
    Ps = np.outer(s, s)
    Gexact = math.exp(-tau / 4) * (np.eye(4) - Ps) + math.exp(3 * tau / 4) * Ps
    assert np.max(np.abs(G - Gexact)) <= 1e-15


# ---------------------------------------------------------------------------
# Network
def test_network_simple_net(ora, golden):
    MA = np.random.default_rng(1).uniform(0, 1, (3, 4))
    MB = np.random.default_rng(2).uniform(0, 1, (4, 3))
    bondsA = [ora.Bond(ora.BD_IN, 3), ora.Bond(ora.BD_OUT, 2), ora.Bond(ora.BD_OUT, 2)]
    A = ora.UniTensor(bondsA)
    B = A.copy()
    A.put_block(MA)
    B.transpose()
    B.put_block(MB)
    net = ora.Network(os.path.join(golden, "Simple.net"))
    net.put_tensor(0, A)
    net.put_tensor(1, B)
    C = net.launch()
    assert np.allclose(C.get_block(), MA @ MB, atol=1e-12)   # testNetwork.cpp:17-37


def test_network_egN1_order_brackets(ora, golden):
    net = ora.Network(os.path.join(golden, "egN1_network"))
    assert net.brakets == [2, 0, 0, -1, 0, -1]
    assert [net.names[i] for i in net.order] == ["W", "H", "WT"]
    H = ora.load_tensor(os.path.join(golden, "egU1_H_U1"))
    W = ora.load_tensor(os.path.join(golden, "egU3_W"))
    WT = ora.load_tensor(os.path.join(golden, "egU3_WT"))
    net.put_tensor("H", H)
    net.put_tensor("W", W)
    net.put_tensor("WT", WT)
    R = net.launch()
    assert net.tree() == (("W", "H"), "WT")
    # dense check: R[-1,-2] = sum W[-1,1,2] H[1,2,3,4] WT[3,4,-2]
    Wd = W.get_raw_elem().reshape(6, 3, 3)
    Hd = H.get_raw_elem().reshape(3, 3, 3, 3)
    WTd = WT.get_raw_elem().reshape(3, 3, 6)
    ref = np.einsum("aij,ijkl,klb->ab", Wd, Hd, WTd)
    assert np.allclose(R.get_raw_elem(), ref, atol=1e-13)


def test_network_parse_errors(ora, tmp_path):
    p = tmp_path / "bad.net"
    p.write_text("A: 1; 2\nB: 2; 3\n")
    with pytest.raises(ValueError):
        ora.Network(str(p))
    p.write_text("A: 1; 2\nB: 2; 3\nTOUT: 1; 3\nORDER: ((A B)\n")
    with pytest.raises(ValueError):
        ora.Network(str(p))
    p.write_text("A: 1; 2\nB: 2; 3\nTOUT: 1; 3")   # unterminated last line is dropped
    with pytest.raises(ValueError):
        ora.Network(str(p))


def test_network_greedy_equals_dense(ora, golden):
    """itebd.net without ORDER: the greedy plan's result equals the dense einsum."""
    chi, d = 5, 2
    ala, blb, la, lb = ora.itebd_init(chi, d, 3)
    b3 = [ora.Bond(ora.BD_IN, chi), ora.Bond(ora.BD_IN, chi), ora.Bond(ora.BD_OUT, d)]
    A = ora.UniTensor(b3, elem=ala.reshape(-1))
    B = ora.UniTensor(b3, elem=blb.reshape(-1))
    U = ora.UniTensor([ora.Bond(ora.BD_IN, d)] * 2 + [ora.Bond(ora.BD_OUT, d)] * 2)
    G = ora.exph(-0.01, ora.heisenberg_two_site())
    U.put_block(G)
    net = ora.Network(os.path.join(golden, "itebd.net"))
    net.put_tensor("ALa", A)
    net.put_tensor("BLb", B)
    net.put_tensor("expH", U)
    C = net.launch()
    assert C.labels == [-1, -2, -3, -4] and C.RBondNum == 1
    ref = np.einsum("acx,cby,xyzw->azbw", ala, blb, G.reshape(2, 2, 2, 2))
    assert np.allclose(C.elem.reshape(ref.shape), ref, atol=1e-14)


# ---------------------------------------------------------------------------
# iTEBD energies
def test_itebd_chi1_product_state_energy_bounds(ora, golden):
    """chi=1: variational energies of a product state lie in [-1/4, 1/4]."""
    t = ora.ITEBD(1, 2, ora.heisenberg_two_site(), 0.05, 9, golden)
    for _ in range(50):
        t.step()
    e = t.energy()
    assert -0.25 - 1e-12 <= e <= 0.25 + 1e-12


def test_itebd_converges_towards_bethe(ora, golden):
    """chi=16, 600 steps at tau=0.05 then 0.01: energy per site approaches
    1/4 - ln 2 from above (variational) within 2e-3."""
    t = ora.ITEBD(16, 2, ora.heisenberg_two_site(), 0.05, 1, golden)
    for _ in range(300):
        t.step()
    t.tau = 0.01
    t.U.put_block(ora.exph(-0.01, t.H))
    for _ in range(300):
        t.step()
    e = t.energy()
    bethe = 0.25 - math.log(2)
    assert bethe - 1e-9 <= e <= bethe + 2e-3


def test_spin1_two_site_closed_form(ora):
    """Two spin-1 Heisenberg (delta = 1): S1.S2 = (S(S+1) - 4) / 2 over total
    spin S = 2, 1, 0, i.e. eigenvalues 1 (x5), -1 (x3), -2 (x1); at delta = 0
    the Sz Sz term drops and the spectrum is symmetric under Sz -> -Sz."""
    ev = np.linalg.eigvalsh(ora.xxz_two_site_spin1(1.0))
    assert np.allclose(ev, [-2.0] + [-1.0] * 3 + [1.0] * 5, atol=1e-13)
    H0 = ora.xxz_two_site_spin1(0.0)
    flip = np.kron(np.eye(3)[::-1], np.eye(3)[::-1])
    assert np.allclose(flip @ H0 @ flip, H0)
    # charge conservation: H only couples |s t> and |s' t'> with s + t = s' + t'
    q = np.add.outer([2, 0, -2], [2, 0, -2]).reshape(-1)
    assert np.all(ora.xxz_two_site_spin1(0.7)[q[:, None] != q[None, :]] == 0)


@pytest.mark.parametrize("d", [2, 3])
def test_u1_sectors_fit_their_theta_blocks(ora, d):
    """Every sector sum equals chi, every sector holds at least one state, and
    no sector holds more states than the theta block that produces it."""
    phys = ora.phys_charges(d)
    for chi in (4, 8, 16, 24, 32, 48, 64, 100, 128, 256, 512, 1024):
        b = ora.u1_bond_sectors(chi, d)
        for p in (0, 1):
            (qR, nR), (qL, nL) = b[p], b[1 - p]
            assert sum(nR) == chi and min(nR) >= 1 and list(qR) == sorted(qR)
            dimL = dict(zip(qL, nL))
            for q, n in zip(qR, nR):
                assert n <= sum(dimL.get(q - s, 0) for s in phys)


@pytest.mark.parametrize("d,chi,delta", [(2, 8, 0.7), (3, 8, 1.3), (2, 16, 1.0)])
def test_u1_update_matches_dense_without_truncation(ora, golden, d, chi, delta):
    """Keeping every singular value, the per-charge-block update and a dense
    update describe the same state: the same starting energy; the singular
    values of one SVD of the whole theta (the dense update of the reference
    example through the oracle's networks when every charge theta reaches is
    a sector of the middle bond, as for d = 2; numpy's SVD of theta with the
    rows and columns of unkept charges zeroed, as the blocks drop them, for
    d = 3); and the rebuilt tensors reproduce the gated theta:
    sum_m X'[a,m,s] Y'[m,b,t] = C[a,s,b,t] / ||S|| on the kept charges."""
    U = ora.ITEBDU1(chi, d, delta, 0.05, 11, golden)
    D = ora.ITEBD(chi, d, ora.xxz_two_site_d(delta, d), 0.05, 1, golden)
    b3 = [ora.Bond(ora.BD_IN, chi), ora.Bond(ora.BD_IN, chi), ora.Bond(ora.BD_OUT, d)]
    D.ALa = ora.UniTensor(b3, "ALa", elem=U.X[0].reshape(-1))
    D.BLb = ora.UniTensor(b3, "BLb", elem=U.X[1].reshape(-1))
    D.La, D.Lb = U.lam[0].copy(), U.lam[1].copy()
    assert abs(U.energy() - D.energy()) <= 1e-13
    X, Y = U.X[0].copy(), U.X[1].copy()
    C = np.einsum("asbt,stuv->aubv", np.einsum("ams,mbt->asbt", X, Y), U.G)
    kept = np.array(U.sect[0][0])
    rowq = (U.q[1][:, None] + np.array(U.phys)[None, :]).reshape(-1)
    colq = (U.q[1][:, None] - np.array(U.phys)[None, :]).reshape(-1)
    rin, cin = np.isin(rowq, kept), np.isin(colq, kept)
    theta = (U.lam[1][:, None, None, None] * C).reshape(chi * d, chi * d) * rin[:, None] * cin[None, :]
    sv = np.linalg.svd(theta, compute_uv=False)
    D.chi = 10 ** 6
    U._update(0, keep_all=True)
    mine = np.sort(U.lam[0])[::-1]
    nrm = np.linalg.norm(sv)
    assert len(mine) <= len(sv)
    assert np.max(np.abs(mine - sv[:len(mine)] / nrm)) <= 1e-13
    assert np.all(sv[len(mine):] <= 1e-13 * nrm)
    if rin.all() and cin.all():
        La = D._update(D.ALa, D.BLb, D.La, D.Lb)[2]
        assert np.max(np.abs(mine - La[:len(mine)])) <= 1e-13
        assert np.all(np.abs(La[len(mine):]) <= 1e-13)
    else:
        assert d == 3
    T = np.einsum("ams,mbt->asbt", U.X[0], U.X[1])
    Ck = C * rin.reshape(chi, d)[:, :, None, None] * cin.reshape(chi, d)[None, None, :, :]
    assert np.max(np.abs(T * nrm - Ck)) <= 1e-13 * np.max(np.abs(C))


def test_u1_truncated_chain_stays_symmetric(ora, golden):
    """With truncation, the tensors keep their charge structure (no entry
    where q_right != q_left + s), the lambdas stay normalised and sorted
    within each sector, and the energy decreases under imaginary time."""
    U = ora.ITEBDU1(16, 2, 1.0, 0.05, 3, golden)
    e0 = U.energy()
    for _ in range(10):
        U.step()
    for p in (0, 1):
        qL, qR = U.q[1 - p], U.q[p]
        mask = qR[None, :, None] == qL[:, None, None] + np.array(U.phys)[None, None, :]
        assert np.all(U.X[p][~mask] == 0)
        assert abs(np.linalg.norm(U.lam[p]) - 1.0) <= 1e-14
        o = 0
        for n in U.sect[p][1]:
            assert np.all(np.diff(U.lam[p][o:o + n]) <= 0)
            o += n
    assert U.energy() < e0


# ---------------------------------------------------------------------------
# Lanczos (lanczosEV, uni10_lapack_cpu.cpp:416-509)
def test_lanczos_four_site_ring_closed_form(ora):
    # four-site periodic Heisenberg ring: ground energy exactly -2 J
    H = ora.heisenberg_ring(4)
    assert np.array_equal(H, H.T)
    e0, psi, it, conv = ora.lanczos_ev(H, np.random.default_rng(1).uniform(-1, 1, 16))
    assert conv and it >= 2
    assert abs(e0 + 2.0) < 1e-13
    assert np.linalg.norm(H @ psi + 2.0 * psi) < 1e-12


@pytest.mark.parametrize("n", [10, 64, 300])
def test_lanczos_lowest_eigenvalue_matches_eigh(ora, n):
    # S = H H^T + H^T H as in the reference's egAP example; the lowest
    # eigenvalue from the textbook dense routine, the vector by its residual
    rng = np.random.default_rng(n)
    X = rng.uniform(-1, 1, (n, n))
    S = X @ X.T + X.T @ X
    start = rng.uniform(-1, 1, n)
    e0, psi, it, conv = ora.lanczos_ev(S, start)
    w = np.linalg.eigvalsh(S)
    assert conv
    assert abs(e0 - w[0]) <= 1e-12 * abs(w[-1])
    assert np.dot(psi, start) >= 0
    assert np.linalg.norm(S @ psi - e0 * psi) <= 1e-4 * abs(w[-1]) * np.linalg.norm(psi)


def test_lanczos_edge_cases(ora):
    with pytest.raises(ValueError):
        ora.lanczos_ev(np.eye(4), np.ones(4), max_iter=2)
    # a start vector that is an eigenvector closes the space after one step:
    # the reference then reports 0 and leaves psi unchanged
    H = np.diag([1.0, 2.0, 3.0])
    e0, psi, it, conv = ora.lanczos_ev(H, np.array([0.0, 1.0, 0.0]))
    assert (e0, it, conv) == (0.0, 1, True)
    assert np.array_equal(psi, [0.0, 1.0, 0.0])
    # max_iter reached before the energy settles: not converged
    rng = np.random.default_rng(7)
    X = rng.uniform(-1, 1, (200, 200))
    e0, psi, it, conv = ora.lanczos_ev(X + X.T, rng.uniform(-1, 1, 200), max_iter=3)
    assert it == 3 and not conv


@pytest.mark.parametrize("n", [10, 64])
def test_lanczos_complex_hermitian_matches_eigh(ora, n):
    # complex form (uni10_lapack_cpu.cpp:905-1000) on S = H H^+ + H^+ H
    rng = np.random.default_rng(50 + n)
    X = rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n))
    S = X @ X.conj().T + X.conj().T @ X
    start = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    e0, psi, it, conv = ora.lanczos_ev(S, start)
    w = np.linalg.eigvalsh(S)
    assert conv
    assert abs(e0 - w[0]) <= 1e-12 * abs(w[-1])
    ov = np.vdot(start, psi)
    assert ov.real > 0 and abs(ov.imag) <= 1e-6 * abs(ov)
    assert np.linalg.norm(S @ psi - e0 * psi) <= 1e-4 * abs(w[-1]) * np.linalg.norm(psi)


# ---------------------------------------------------------------------------
# Greedy contraction tree: the known answer printed in Network.h:145-193
def test_greedy_tree_matches_network_h_known_answer(ora, tmp_path):
    p = tmp_path / "network_h.net"
    p.write_text(NETWORK_H_NET)
    net = ora.Network(str(p))
    for nm, labs, r in zip(net.names[:-1], net.label_arr[:-1], net.Rnums[:-1]):
        T = ora.UniTensor(network_h_bonds(ora, labs, r), nm)
        net.put_tensor(nm, T)
    assert net.print_tree() == NETWORK_H_PRINT


# ---------------------------------------------------------------------------
# Fermionic signs.  The convention (UniTensorReal.cpp:657-681): a tensor's
# bonds are ordered row bonds first, then column bonds in reverse, and a
# permute reorders them as Grassmann generators, one per fermion-odd index.
# The checks below compute that sign independently (cycle decomposition of
# the reordering of the odd generators), and check the algebra it must obey.
def _grassmann_sign_mask(T, rsp, rb_out):
    """Per element of the dense tensor: the sign of reordering the odd
    generators from the input order (rows, then reversed columns) to the
    output order, as the parity of the cycle decomposition."""
    nb = len(T.bonds)
    rb = T.RBondNum
    fin = list(range(rb)) + list(range(nb - 1, rb - 1, -1))
    fout = [rsp[j] for j in range(rb_out)] + [rsp[j] for j in range(nb - 1, rb_out - 1, -1)]
    par = [np.array([q.prtF for q in b.Qlist()]) for b in T.bonds]
    dims = [b.dim for b in T.bonds]
    mask = np.ones(dims)
    for idx in np.ndindex(*dims):
        odd = {b for b in range(nb) if par[b][idx[b]]}
        a = [b for b in fin if b in odd]
        t = [b for b in fout if b in odd]
        pos = {b: k for k, b in enumerate(a)}
        order = [pos[b] for b in t]
        mask[idx] = _perm_sign_by_cycles(order)
    return mask


@pytest.mark.parametrize("seed", range(12))
def test_fermionic_permute_is_the_grassmann_reordering(ora, monkeypatch, seed):
    monkeypatch.setattr(ora.Qnum, "fermionic", True)
    rng = np.random.default_rng(300 + seed)
    nb = int(rng.integers(2, 5))
    T = fermi_tensor(ora, rng, nb, int(rng.integers(0, nb + 1)))
    if T is None:
        pytest.skip("no charge-conserving layout drawn")
    perm = [int(v) for v in rng.permutation(nb)]
    rb_out = int(rng.integers(0, nb + 1))
    dense = T.get_raw_elem().reshape([b.dim for b in T.bonds])
    try:
        P = T.copy().permute([T.labels[p] for p in perm], rb_out)
    except ValueError:
        pytest.skip("permuted bonds carry no common charge")
    want = np.transpose(dense * _grassmann_sign_mask(T, perm, rb_out), perm)
    got = P.get_raw_elem().reshape([b.dim for b in P.bonds])
    assert np.array_equal(got, want)


@pytest.mark.parametrize("seed", range(12))
def test_fermionic_permute_round_trip_and_composition(ora, monkeypatch, seed):
    """A permute followed by its inverse is the identity bit for bit, and two
    permutes (each with its own row-bond count) equal one permute to the
    final order: the signs compose as the reorderings do."""
    monkeypatch.setattr(ora.Qnum, "fermionic", True)
    rng = np.random.default_rng(400 + seed)
    nb = int(rng.integers(2, 5))
    T = fermi_tensor(ora, rng, nb, int(rng.integers(0, nb + 1)))
    if T is None:
        pytest.skip("no charge-conserving layout drawn")
    l1 = [T.labels[p] for p in rng.permutation(nb)]
    l2 = [T.labels[p] for p in rng.permutation(nb)]
    r1, r2 = int(rng.integers(0, nb + 1)), int(rng.integers(0, nb + 1))
    try:
        two = T.copy().permute(l1, r1).permute(l2, r2)
        one = T.copy().permute(l2, r2)
        back = T.copy().permute(l1, r1).permute(T.labels, T.RBondNum)
    except ValueError:
        pytest.skip("an intermediate layout carries no common charge")
    assert two.elem.tobytes() == one.elem.tobytes()
    assert back.elem.tobytes() == T.elem.tobytes()


def test_fermionic_sign_known_answers(ora, monkeypatch):
    """Rank 2, both bonds carrying one even and one odd state: exchanging
    the two row bonds flips exactly the odd-odd element; moving the column
    bond into the rows keeps every sign (rows then reversed columns is the
    same generator order), and moving the row bond into the columns
    reverses the two (odd-odd flips)."""
    monkeypatch.setattr(ora.Qnum, "fermionic", True)
    Q = ora.Qnum
    qs = [Q(0, ora.PRT_EVEN, ora.PRTF_EVEN), Q(0, ora.PRT_EVEN, ora.PRTF_ODD)]

    def make(rb):
        T = ora.UniTensor([ora.Bond(ora.BD_IN if b < rb else ora.BD_OUT, qs) for b in range(2)])
        T.set_label([1, 2])
        return T
    raw = np.array([[1.0, 0.0], [0.0, 4.0]])        # parity conserved: (e,e), (o,o)
    T = make(2)
    T.set_raw_elem(raw)
    assert np.array_equal(T.copy().permute([2, 1], 2).get_raw_elem().reshape(2, 2), [[1.0, 0.0], [0.0, -4.0]])
    T1 = make(1)
    T1.set_raw_elem(raw)
    assert np.array_equal(T1.copy().permute([1, 2], 2).get_raw_elem().reshape(2, 2), raw)
    assert np.array_equal(T1.copy().permute([1, 2], 0).get_raw_elem().reshape(2, 2), [[1.0, 0.0], [0.0, -4.0]])


def test_add_gate_is_the_elementwise_parity_sign(ora, monkeypatch):
    """addGate multiplies element (i_0, .., i_n) by (-1)^(sum over gates of
    p(i_b1) p(i_b2)); computed here per dense element from its indices."""
    monkeypatch.setattr(ora.Qnum, "fermionic", True)
    rng = np.random.default_rng(7)
    done = 0
    for _ in range(40):
        nb = int(rng.integers(2, 5))
        T = fermi_tensor(ora, rng, nb, int(rng.integers(0, nb + 1)))
        if T is None:
            continue
        gates = [tuple(int(v) for v in rng.choice(nb, 2, replace=False)) for _ in range(int(rng.integers(1, 4)))]
        dense = T.get_raw_elem().reshape([b.dim for b in T.bonds])
        par = [np.array([q.prtF for q in b.Qlist()]) for b in T.bonds]
        sign = np.ones(dense.shape)
        for idx in np.ndindex(*dense.shape):
            s = sum(int(par[a][idx[a]]) * int(par[b][idx[b]]) for a, b in gates)
            sign[idx] = -1.0 if s % 2 else 1.0
        G = T.copy().add_gate(gates)
        assert np.array_equal(G.get_raw_elem().reshape(dense.shape), dense * sign)
        done += 1
    assert done >= 20


def test_add_swap_known_answer(ora, monkeypatch, tmp_path):
    """Network ((A C) B) over the chain A-B-C: the tree lists the leaves as
    A, C, B, so B passes C once; B's bond shared with C (label 3, position 1)
    crosses its other bond (label 2, position 0): one gate (0, 1) on B and
    none elsewhere.  That gate cancels the signs the permutes inside the
    contractions pick up along this order, so the result is the one of the
    order ((A B) C), which needs no gate and no sign: the plain product of
    the three matrices (fermionic contraction does not depend on the order)."""
    monkeypatch.setattr(ora.Qnum, "fermionic", True)
    Q = ora.Qnum
    qs = [Q(0, ora.PRT_EVEN, ora.PRTF_EVEN), Q(0, ora.PRT_EVEN, ora.PRTF_ODD), Q(0, ora.PRT_EVEN, ora.PRTF_ODD)]
    rng = np.random.default_rng(5)
    ts = {}
    for nm in "ABC":
        T = ora.UniTensor([ora.Bond(ora.BD_IN, qs), ora.Bond(ora.BD_OUT, qs)], nm)
        T.set_elem(rng.uniform(-1, 1, T.elem_num()))
        ts[nm] = T
    d = {nm: ts[nm].get_raw_elem() for nm in "ABC"}
    want = d["A"] @ d["B"] @ d["C"]
    for order, gates in (("((A C) B)", [[], [(0, 1)], []]), ("((A B) C)", [[], [], []])):
        p = tmp_path / "chain3.net"
        p.write_text(CHAIN3_NET % order)
        net = ora.Network(str(p))
        for nm in "ABC":
            net.put_tensor(nm, ts[nm])
        net.construct()
        assert net.swaps_arr == gates
        assert np.allclose(net.launch().get_raw_elem(), want, rtol=0, atol=1e-14)
        # a second launch does not apply the gates again
        assert np.allclose(net.launch().get_raw_elem(), want, rtol=0, atol=1e-14)


def test_fermionic_network_result_independent_of_order(ora, monkeypatch, tmp_path):
    """Fermionic contraction of a closed network is one number whatever the
    tree: the greedy tree and two bracketed trees (each with its own swap
    gates) must give the same scalar."""
    monkeypatch.setattr(ora.Qnum, "fermionic", True)
    vals, gatesets = [], []
    for k, order in enumerate(FERMI_ORDERS):
        p = tmp_path / ("f%d.net" % k)
        p.write_text(NETWORK_H_NET.replace("ORDER: " + FERMI_ORDERS[0], "ORDER: " + order))
        net = ora.Network(str(p))
        rng = np.random.default_rng(1)
        for nm, labs, r in zip(net.names[:-1], net.label_arr[:-1], net.Rnums[:-1]):
            T = ora.UniTensor(network_h_fermionic_bonds(ora, labs, r), nm)
            T.set_elem(rng.uniform(-1, 1, T.elem_num()))
            net.put_tensor(nm, T)
        vals.append(float(net.launch().elem[0]))
        gatesets.append(net.swaps_arr)
    assert gatesets[0] != gatesets[2]                  # the trees need different gates
    assert max(vals) - min(vals) <= 1e-13 * max(1.0, abs(vals[0]))


def _dense_of(v, shape, diag):
    if not diag:
        return np.asarray(v).reshape(shape)
    D = np.zeros(shape)
    k = min(shape)
    D[np.arange(k), np.arange(k)] = np.asarray(v)[:k]
    return D


RDOTR_CASES = [((5, 7), False, (7, 4), False), ((6, 6), True, (6, 3), False), ((4, 6), True, (6, 5), False),
               ((5, 6), False, (6, 6), True), ((5, 6), False, (6, 4), True), ((6, 6), True, (6, 6), True),
               ((3, 5), True, (5, 4), True)]


@pytest.mark.parametrize("case", RDOTR_CASES)
def test_rdotr_is_the_product_of_the_dense_matrices(ora, case):
    """Every branch of RDotR (dense/diagonal operands, square and
    rectangular diagonal matrices) equals numpy's product of the operands
    written out densely."""
    sa, da, sb, db = case
    rng = np.random.default_rng(sum(sa) + 10 * sum(sb))
    a = rng.uniform(-1, 1, min(sa) if da else sa[0] * sa[1])
    b = rng.uniform(-1, 1, min(sb) if db else sb[0] * sb[1])
    c, shape, diag = ora.rdotr(a, sa, da, b, sb, db)
    assert shape == (sa[0], sb[1]) and diag == (da and db)
    want = _dense_of(a, sa, da) @ _dense_of(b, sb, db)
    assert np.allclose(_dense_of(c, shape, diag), want, rtol=0, atol=1e-15)


def test_otimes_is_the_outer_product(ora):
    """otimes of two U(1) tensors: the dense result is the outer product with
    bonds ordered (A rows, B rows, A columns, B columns)."""
    A = _random_u1(ora, 4)
    B = _random_u1(ora, 5, n_in=1, n_out=2)
    O = ora.otimes(A, B)
    da = A.get_raw_elem().reshape([b.dim for b in A.bonds])
    db = B.get_raw_elem().reshape([b.dim for b in B.bonds])
    ra, rb = A.RBondNum, B.RBondNum
    na, nb = len(A.bonds), len(B.bonds)
    outer = np.multiply.outer(da, db)
    order = list(range(ra)) + [na + i for i in range(rb)] + list(range(ra, na)) + [na + i for i in range(rb, nb)]
    want = np.transpose(outer, order)
    assert O.RBondNum == ra + rb
    assert np.array_equal(O.get_raw_elem().reshape(want.shape), want)
