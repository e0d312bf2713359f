"""GPU parity of the drop-in Uni10 API (C++ over HIP) against the oracle.

Bars: block layout, setRawElem/getRawElem, permute, transpose and file I/O
are bit-exact; contractions are within 1e-12 relative Frobenius error of the
oracle's dgemm; network plans (contraction trees) are identical; iTEBD
energies and Schmidt spectra agree with the oracle's to the stated bounds.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def u10(tn):
    from trim_amd import uni10
    uni10.load()
    uni10.set_gemm_mode(uni10.GEMM_F64)
    return uni10


def to_product(u10, T):
    """Product UniTensor with the oracle tensor's bonds, labels and elements."""
    bonds = [u10.Bond(b.type, [q.triple() for q in b.Qlist()]) for b in T.bonds]
    P = u10.UniTensor(bonds, T.name)
    P.setLabel(T.labels)
    P.setElem(T.elem)
    return P


def same_layout(P, T):
    pb = [(q.triple(), r, c, o) for q, r, c, o in P.blocks()]
    tb = [(q.triple(), r, c, o) for q, (r, c, o) in T.layout.blocks.items()]
    return pb == tb


HEIS_S1 = np.array([
    1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, -1, 0, 1, 0, 0, 0, 0,
    0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0,
    0, 0, 0, 0, 1, 0, -1, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1], float)


def test_grouping_setraw_matches_golden(u10, ora, golden):
    qs = [(1, 0, 0), (0, 0, 0), (-1, 0, 0)]
    H = u10.UniTensor([u10.Bond(1, qs), u10.Bond(1, qs), u10.Bond(-1, qs), u10.Bond(-1, qs)], "H_U1")
    H.setRawElem(HEIS_S1)
    G = ora.load_tensor(os.path.join(golden, "egU1_H_U1"))
    assert H.getElem().tobytes() == G.elem.tobytes()
    assert same_layout(H, G)
    assert np.array_equal(H.getRawElem().reshape(-1), HEIS_S1)
    assert H.at([0, 2, 1, 1]) == HEIS_S1.reshape(3, 3, 3, 3)[0, 2, 1, 1]


def test_file_loader_and_transpose_golden(u10, ora, golden, tmp_path):
    W = u10.UniTensor.load(os.path.join(golden, "egU3_W"))
    WT = ora.load_tensor(os.path.join(golden, "egU3_WT"))
    Wo = ora.load_tensor(os.path.join(golden, "egU3_W"))
    assert W.getElem().tobytes() == Wo.elem.tobytes()
    assert W.label() == [-1, 1, 2]
    W.transpose()
    assert W.getElem().tobytes() == WT.elem.tobytes()
    assert same_layout(W, WT)
    # save (current format) and read back
    f = str(tmp_path / "w.bin")
    W.save(f)
    R = u10.UniTensor.load(f)
    assert R.getElem().tobytes() == W.getElem().tobytes() and R.label() == W.label()
    assert ora.load_tensor(f).elem.tobytes() == W.getElem().tobytes()


def _random_u1(ora, seed, n_in=2, n_out=2, maxdeg=4):
    rng = np.random.default_rng(seed)
    Q = ora.Qnum
    for attempt in range(50):
        bonds = []
        for i in range(n_in + n_out):
            tp = ora.BD_IN if i < n_in else ora.BD_OUT
            charges = sorted(set(int(c) for c in rng.integers(-3, 4, size=int(rng.integers(1, 4)))))
            qs = []
            for c in charges:
                qs += [Q(c)] * int(rng.integers(1, maxdeg + 1))
            if rng.random() < 0.3:
                qs += [qs[0]]
            bonds.append(ora.Bond(tp, qs))
        try:
            T = ora.UniTensor(bonds)
        except ValueError:
            continue
        T.set_raw_elem(rng.uniform(-1, 1, size=[b.dim for b in bonds]))
        return T
    raise RuntimeError("no valid tensor")


@pytest.mark.parametrize("seed", range(16))
def test_symmetric_permute_bitexact(u10, ora, seed):
    rng = np.random.default_rng(500 + seed)
    T = _random_u1(ora, seed, n_in=int(rng.integers(1, 3)), n_out=int(rng.integers(1, 3)))
    nb = len(T.bonds)
    perm = [int(v) for v in rng.permutation(nb)]
    rb = int(rng.integers(0, nb + 1))
    labels = [T.labels[p] for p in perm]
    _say("sym permute seed=%d nb=%d perm=%s rb=%d" % (seed, nb, perm, rb))
    P = to_product(u10, T)
    assert same_layout(P, T)
    try:
        To = T.copy().permute(labels, rb)
    except ValueError:
        with pytest.raises(u10.Uni10Error):
            P.permute(labels, rb)
        return
    P.permute(labels, rb)
    assert P.label() == labels and P.inBondNum() == rb
    assert same_layout(P, To)
    assert P.getElem().tobytes() == To.elem.tobytes()


def test_dense_permute_workload_shape(u10, ora):
    dims = (256, 2, 2, 256)
    bonds = [ora.Bond(ora.BD_IN, 256), ora.Bond(ora.BD_IN, 2), ora.Bond(ora.BD_OUT, 2), ora.Bond(ora.BD_OUT, 256)]
    T = ora.UniTensor(bonds, elem=ora.splitmix64_uniform(9, int(np.prod(dims))))
    T.set_label([0, 1, 4, 5])
    P = to_product(u10, T)
    To = T.copy().permute([5, 1, 0, 4], 2)
    P.permute([5, 1, 0, 4], 2)
    assert P.getElem().tobytes() == To.elem.tobytes()


def _pair_u1(ora, seed):
    rng = np.random.default_rng(seed)
    Q = ora.Qnum
    def bond(tp, charges):
        qs = []
        for c in charges:
            qs += [Q(c)] * int(rng.integers(1, 6))
        return ora.Bond(tp, qs)
    chi_c = [-3, -1, 1, 3]
    mid = [-4, -2, 0, 2, 4]
    a_l = bond(ora.BD_IN, chi_c)
    p = ora.Bond(ora.BD_OUT, [Q(1), Q(-1)])
    m_out = bond(ora.BD_OUT, mid)
    A = ora.UniTensor([a_l, p, m_out])
    A.set_raw_elem(rng.uniform(-1, 1, [b.dim for b in A.bonds]))
    A.set_label([1, 2, 3])
    m_in = m_out.copy()
    m_in.type = ora.BD_IN
    B = ora.UniTensor([m_in, ora.Bond(ora.BD_OUT, [Q(1), Q(-1)]), bond(ora.BD_OUT, chi_c)])
    B.set_raw_elem(rng.uniform(-1, 1, [b.dim for b in B.bonds]))
    B.set_label([3, 4, 5])
    return A, B


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("fast", [False, True])
def test_symmetric_contract(u10, ora, seed, fast):
    A, B = _pair_u1(ora, seed)
    PA, PB = to_product(u10, A), to_product(u10, B)
    C = ora.contract(A, B, fast)
    PC = u10.contract(PA, PB, fast)
    assert PC.label() == C.labels and same_layout(PC, C)
    ref = C.elem
    err = np.linalg.norm(PC.getElem() - ref) / np.linalg.norm(ref)
    assert err <= 1e-12
    # inputs: restored when fast is False, left permuted when True (like the reference)
    assert PA.label() == A.labels and PB.label() == B.labels
    assert PA.getElem().tobytes() == A.elem.tobytes()


def test_contract_dense_workload_full_size(u10, ora):
    """A(256,2,2,256) x B(256,2,2,256) over two bonds, then permute (3,1,0,2)."""
    chi = 256
    mk = lambda: [ora.Bond(ora.BD_IN, chi), ora.Bond(ora.BD_IN, 2), ora.Bond(ora.BD_OUT, 2), ora.Bond(ora.BD_OUT, chi)]
    n = chi * 4 * chi
    A = ora.UniTensor(mk(), elem=ora.splitmix64_uniform(1, n))
    B = ora.UniTensor(mk(), elem=ora.splitmix64_uniform(2, n))
    A.set_label([0, 1, 2, 3])
    B.set_label([3, 2, 4, 5])
    PA, PB = to_product(u10, A), to_product(u10, B)
    C = ora.contract(A, B)
    PC = u10.contract(PA, PB)
    assert PC.label() == [0, 1, 4, 5]
    err = np.linalg.norm(PC.getElem() - C.elem) / np.linalg.norm(C.elem)
    assert err <= 1e-12, err
    # the permute is bit-exact: the oracle's permute of the product's own contraction output
    Cc = C.copy()
    Cc.set_elem(PC.getElem())
    Cc.permute([5, 1, 0, 4], 2)
    PC.permute([5, 1, 0, 4], 2)
    assert PC.label() == [5, 1, 0, 4]
    assert PC.getElem().tobytes() == Cc.elem.tobytes()


def test_scalar_contraction_and_otimes(u10, ora):
    A, B = _pair_u1(ora, 3)
    PA, PB = to_product(u10, A), to_product(u10, B)
    # full contraction to a scalar: A with itself
    A2 = A.copy()
    P2 = to_product(u10, A)
    s = ora.contract(A, A2)
    ps = u10.contract(PA, P2)
    assert ps.bondNum() == 0
    assert abs(ps.getElem()[0] - s.elem[0]) <= 1e-12 * abs(s.elem[0])
    o = u10.otimes(PA, PB)
    assert o.bondNum() == 6


def test_network_simple_and_order(u10, ora, golden):
    # Simple.net (testNetwork.cpp:17-37)
    MA = np.random.default_rng(1).uniform(0, 1, (3, 4))
    MB = np.random.default_rng(2).uniform(0, 1, (4, 3))
    A = u10.UniTensor([u10.Bond(1, 3), u10.Bond(-1, 2), u10.Bond(-1, 2)])
    A.putBlock(MA)
    B = A.copy()
    B.transpose()
    B.putBlock(MB)
    net = u10.Network(os.path.join(golden, "Simple.net"))
    net.putTensor("A", A)
    net.putTensor("B", B)
    C = net.launch()
    assert np.max(np.abs(C.getBlock() - MA @ MB)) <= 1e-12
    # egN1_network with ORDER brackets on the golden U(1) tensors
    net = u10.Network(os.path.join(golden, "egN1_network"))
    H = u10.UniTensor.load(os.path.join(golden, "egU1_H_U1"))
    W = u10.UniTensor.load(os.path.join(golden, "egU3_W"))
    WT = u10.UniTensor.load(os.path.join(golden, "egU3_WT"))
    for nm, t in (("H", H), ("W", W), ("WT", WT)):
        net.putTensor(nm, t)
    R = net.launch()
    assert net.tree() == "((W H) WT)"
    onet = ora.Network(os.path.join(golden, "egN1_network"))
    for nm in ("H", "W", "WT"):
        onet.put_tensor(nm, ora.load_tensor(os.path.join(golden, "eg%s" % {"H": "U1_H_U1", "W": "U3_W", "WT": "U3_WT"}[nm])))
    Ro = onet.launch()
    assert np.max(np.abs(R.getElem() - Ro.elem)) <= 1e-12


def _otree(t):
    return t if isinstance(t, str) else "(%s %s)" % (_otree(t[0]), _otree(t[1]))


def test_network_greedy_plan_matches_oracle(u10, ora, golden):
    chi, d = 6, 2
    ala, blb, la, lb = ora.itebd_init(chi, d, 3)
    b3 = [ora.Bond(ora.BD_IN, chi), ora.Bond(ora.BD_IN, chi), ora.Bond(ora.BD_OUT, d)]
    A = ora.UniTensor(b3, elem=ala.reshape(-1))
    B = ora.UniTensor(b3, elem=blb.reshape(-1))
    U = ora.UniTensor([ora.Bond(ora.BD_IN, d)] * 2 + [ora.Bond(ora.BD_OUT, d)] * 2)
    U.put_block(ora.exph(-0.01, ora.heisenberg_two_site()))
    onet = ora.Network(os.path.join(golden, "itebd.net"))
    pnet = u10.Network(os.path.join(golden, "itebd.net"))
    for nm, t in (("ALa", A), ("BLb", B), ("expH", U)):
        t.name = nm
        onet.put_tensor(nm, t)
        pnet.putTensor(nm, to_product(u10, t))
    Co = onet.launch()
    Cp = pnet.launch()
    assert pnet.tree() == _otree(onet.tree())
    assert Cp.label() == Co.labels
    assert np.max(np.abs(Cp.getElem() - Co.elem)) <= 1e-14
    # a second launch reuses the plan (leaves stay permuted, as in the reference)
    Cp2 = pnet.launch()
    assert np.max(np.abs(Cp2.getElem() - Co.elem)) <= 1e-14


MODE_TOL = {0: 1e-12, 1: 2e-6, 2: 5e-5, 3: 2e-6, 4: 1e-11}


@pytest.mark.parametrize("size", [(96, 4, 21), (1024, 8, 11)], ids=["chi96", "chi1024"])
@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_u1_two_site_contraction_modes(u10, ora, mode, size):
    """Theta = (A1 . A2) . G for U(1) tensors with binomial sectors, at chi=96
    over Sz in {-4..4} and at the configured chi=1024 over Sz in {-8..8}:
    FP64 within 1e-12 of the oracle; FP32-MFMA and split-BF16 within their
    stated tolerances."""
    from trim_amd import workloads
    chi, qmax, seed = size
    _say("U(1) two-site contraction chi=%d qmax=%d mode=%d" % (chi, qmax, mode))
    A1, A2, G, dims = workloads.u1_two_site(u10, chi, qmax, seed, 0.01)
    Q = ora.Qnum
    charges = list(range(-qmax, qmax + 1))
    def ob(tp, ch, ds):
        qs = []
        for c, d in zip(ch, ds):
            qs += [Q(c)] * d
        return ora.Bond(tp, qs)
    ci, co = ob(ora.BD_IN, charges, dims), ob(ora.BD_OUT, charges, dims)
    po, pi = ob(ora.BD_OUT, [1, -1], [1, 1]), ob(ora.BD_IN, [1, -1], [1, 1])
    oA1 = ora.UniTensor([ci, po, co], elem=workloads.splitmix_uniform(seed, A1.elemNum()))
    oA2 = ora.UniTensor([ci, po, co], elem=workloads.splitmix_uniform(seed + 1, A2.elemNum()))
    oA1.set_label([-1, 10, 0])
    oA2.set_label([0, 11, -2])
    oG = ora.UniTensor([pi, pi, po, po])
    oG.set_raw_elem(ora.exph(-0.01, workloads.sz_hamiltonian(seed + 2)))
    oG.set_label([10, 11, 20, 21])
    assert same_layout(A1, oA1) and same_layout(G, oG)
    ref = ora.contract(ora.contract(oA1, oA2), oG)
    u10.set_gemm_mode(mode)
    try:
        R = u10.contract(u10.contract(A1, A2), G)
    finally:
        u10.set_gemm_mode(0)
    assert R.label() == ref.labels and same_layout(R, ref)
    err = np.linalg.norm(R.getElem() - ref.elem) / np.linalg.norm(ref.elem)
    _say("u1 two-site mode %d rel err %.3e" % (mode, err))
    assert err <= MODE_TOL[mode]


def test_svd_eigh_exph_api(u10, ora):
    a = np.random.default_rng(4).uniform(-1, 1, (12, 7))
    U, S, VT = u10.svd(a)
    assert np.allclose(U @ np.diag(S) @ VT, a, atol=1e-13)
    assert np.allclose(S, ora.svd(a)[1], rtol=1e-13, atol=0)
    H = ora.heisenberg_two_site(0.7)
    G = u10.exph(-0.05, H)
    assert np.max(np.abs(G - ora.exph(-0.05, H))) <= 1e-15


def test_itebd_matches_oracle_short(u10, ora, golden):
    """chi=8, 10 steps: energies and Schmidt spectra agree with the oracle."""
    chi, tau, seed = 8, 0.01, 11
    P = u10.ITebd(chi, 2, 1.0, tau, seed, golden)
    O = ora.ITEBD(chi, 2, ora.heisenberg_two_site(1.0), tau, seed, golden)
    e0p, e0o = P.energy(), O.energy()
    assert abs(e0p - e0o) <= 1e-12, (e0p, e0o)
    for k in range(10):
        P.step()
        O.step()
    ep, eo = P.energy(), O.energy()
    _say("itebd chi=8: product %.15f oracle %.15f" % (ep, eo))
    assert abs(ep - eo) <= 1e-10
    assert np.max(np.abs(P.lam(0) - O.La)) <= 1e-10
    assert np.max(np.abs(P.lam(1) - O.Lb)) <= 1e-10


def test_itebd_thousand_steps_matches_oracle(u10, ora, golden):
    """The full single-chain workload: chi=32 spin-1/2 Heisenberg, tau=0.01,
    1000 steps through the drop-in API (warm-started, QR-preconditioned SVDs
    chained over 2000 updates) against the oracle's iTEBD: energy per site to
    1e-12 and both Schmidt spectra to 1e-10 (measured: 2e-14 and 5e-12)."""
    chi, tau, seed = 32, 0.01, 7
    P = u10.ITebd(chi, 2, 1.0, tau, seed, golden)
    O = ora.ITEBD(chi, 2, ora.heisenberg_two_site(1.0), tau, seed, golden)
    for _ in range(1000):
        P.step()
        O.step()
    ep, eo = P.energy(), O.energy()
    _say("itebd chi=32 after 1000 steps: %.15f oracle %.15f" % (ep, eo))
    assert abs(ep - eo) <= 1e-12
    assert np.max(np.abs(P.lam(0) - O.La)) <= 1e-10
    assert np.max(np.abs(P.lam(1) - O.Lb)) <= 1e-10


# ---------------------------------------------------------------------------
# batched iTEBD (many chains per launch)
BATCH = dict(chi=[4, 8, 12, 40], delta=[0.5, 1.0, 1.5, 1.0], seed=[11, 12, 13, 14])


def test_batch_tebd_matches_oracle_per_chain(u10, ora, golden):
    """Every chain of a batch of uneven bond dimensions (small-kernel and
    block-Jacobi SVD paths) follows the oracle's single-chain iTEBD: energies
    and Schmidt spectra to 1e-10 after 6 steps."""
    tau = 0.01
    B = u10.BatchTebd(BATCH["chi"], BATCH["delta"], BATCH["seed"], tau, golden)
    B.step(6)
    for i, (chi, delta, seed) in enumerate(zip(BATCH["chi"], BATCH["delta"], BATCH["seed"])):
        O = ora.ITEBD(chi, 2, ora.heisenberg_two_site(delta), tau, seed, golden)
        for _ in range(6):
            O.step()
        e, eo = B.energy(i), O.energy()
        _say("batch chain %d chi=%d delta=%.1f: %.15f oracle %.15f" % (i, chi, delta, e, eo))
        assert abs(e - eo) <= 1e-10
        assert np.max(np.abs(B.lam(i, 0) - O.La)) <= 1e-10
        assert np.max(np.abs(B.lam(i, 1) - O.Lb)) <= 1e-10
        assert abs(B.trunc_err(i) - O.trunc_err) <= 1e-10


def test_batch_tebd_graph_replay_is_identical(u10, golden):
    """Replaying the captured step gives the same bits as issuing it."""
    a = u10.BatchTebd(BATCH["chi"], BATCH["delta"], BATCH["seed"], 0.01, golden)
    b = u10.BatchTebd(BATCH["chi"], BATCH["delta"], BATCH["seed"], 0.01, golden)
    b.set_graph(True)
    for _ in range(3):
        a.step(2)
        b.step(2)
    for i in range(len(BATCH["chi"])):
        for w in (0, 1):
            assert a.lam(i, w).tobytes() == b.lam(i, w).tobytes()


def test_batch_tebd_fp32_mode_error_is_stated(u10, golden):
    """FP32-MFMA contractions (SVD and truncation FP64): energies within
    1e-6 of the FP64 batch after 10 steps."""
    a = u10.BatchTebd(BATCH["chi"], BATCH["delta"], BATCH["seed"], 0.01, golden, gemm_mode=0)
    b = u10.BatchTebd(BATCH["chi"], BATCH["delta"], BATCH["seed"], 0.01, golden, gemm_mode=1)
    a.step(10)
    b.step(10)
    for i in range(len(BATCH["chi"])):
        assert abs(a.energy(i) - b.energy(i)) <= 1e-6


MIXED = dict(chi=[8, 16, 8, 8, 12, 40, 24], delta=[0.7, 1.0, 1.3, 0.5, 1.0, 0.9, 0.4],
             seed=[31, 32, 33, 34, 35, 36, 37], d=[2, 2, 3, 3, 2, 2, 3], sym=[1, 1, 1, 0, 0, 1, 1])


def _mixed(u10, golden, tau, **kw):
    return u10.BatchTebd(MIXED["chi"], MIXED["delta"], MIXED["seed"], tau, golden, d=MIXED["d"],
                         sym=MIXED["sym"], **kw)


def test_batch_tebd_mixed_kinds_match_oracle(u10, ora, golden):
    """One batch of dense and U(1) chains, spin 1/2 and spin 1: every chain
    follows its oracle (the dense iTEBD of the reference example, or the
    per-charge-block U(1) update): sector tables, energies, Schmidt spectra
    (sector by sector for U(1)) and discarded weight to 1e-10 after 5 steps."""
    tau = 0.02
    B = _mixed(u10, golden, tau)
    B.step(5)
    for i, (chi, delta, seed, d, sym) in enumerate(zip(*(MIXED[k] for k in ("chi", "delta", "seed", "d", "sym")))):
        if sym:
            O = ora.ITEBDU1(chi, d, delta, tau, seed, golden)
            for b in (0, 1):
                q, n = B.sectors(i, b)
                assert list(q) == list(O.sect[b][0]) and list(n) == list(O.sect[b][1])
        else:
            O = ora.ITEBD(chi, d, ora.xxz_two_site_d(delta, d), tau, seed, golden)
            assert len(B.sectors(i, 0)[0]) == 0
        for _ in range(5):
            O.step()
        la, lb = (O.lam[0], O.lam[1]) if sym else (O.La, O.Lb)
        e, eo = B.energy(i), O.energy()
        _say("mixed chain %d chi=%d d=%d %s: %.15f oracle %.15f" % (i, chi, d, "U(1)" if sym else "dense", e, eo))
        assert abs(e - eo) <= 1e-10
        assert np.max(np.abs(B.lam(i, 0) - la)) <= 1e-10
        assert np.max(np.abs(B.lam(i, 1) - lb)) <= 1e-10
        assert abs(B.trunc_err(i) - O.trunc_err) <= 1e-10


@pytest.mark.parametrize("chi", [300, 520])
def test_batch_tebd_large_dense_chain_matches_oracle(u10, ora, golden, chi):
    """A dense spin-1/2 chain whose theta (2 chi square) goes through the
    block SVD with QR preconditioning: 32-column panels at chi = 300 (600
    rows), 16-column ones at chi = 520 (1040 rows).  Two steps against the
    oracle's dense iTEBD: energy, both Schmidt spectra and the discarded
    weight to 1e-10."""
    tau, delta, seed = 0.02, 0.8, 41
    _say("large dense chain chi=%d" % chi)
    B = u10.BatchTebd([chi], [delta], [seed], tau, golden)
    B.step(2)
    O = ora.ITEBD(chi, 2, ora.xxz_two_site_d(delta, 2), tau, seed, golden)
    for _ in range(2):
        O.step()
    assert abs(B.energy(0) - O.energy()) <= 1e-10
    assert np.max(np.abs(B.lam(0, 0) - O.La)) <= 1e-10
    assert np.max(np.abs(B.lam(0, 1) - O.Lb)) <= 1e-10
    assert abs(B.trunc_err(0) - O.trunc_err) <= 1e-10


def test_batch_tebd_mixed_graph_replay_and_slot_reload(u10, golden):
    """The mixed batch replayed as a graph gives the same bits as issued
    step by step; reloading one U(1) slot restarts it exactly (same bits as
    a fresh engine after the same steps) without touching the others."""
    a = _mixed(u10, golden, 0.02)
    b = _mixed(u10, golden, 0.02)
    b.set_graph(True)
    for _ in range(2):
        a.step(2)
        b.step(2)
    for i in range(len(MIXED["chi"])):
        for w in (0, 1):
            assert a.lam(i, w).tobytes() == b.lam(i, w).tobytes()
    keep = [b.lam(i, 0).tobytes() for i in range(len(MIXED["chi"]))]
    b.load(1, MIXED["delta"][1], MIXED["seed"][1])
    b.step(3)
    c = _mixed(u10, golden, 0.02)
    c.set_graph(True)
    c.step(3)
    assert b.lam(1, 0).tobytes() == c.lam(1, 0).tobytes()
    assert b.lam(1, 1).tobytes() == c.lam(1, 1).tobytes()
    assert b.lam(0, 0).tobytes() != keep[0]   # the others kept evolving


# ---------------------------------------------------------------------------
# lanczosEigh (Block.cpp:474-506 -> lanczosEV) on the device
@pytest.mark.parametrize("n", [10, 64, 300, 1000])
def test_lanczos_matches_oracle(u10, ora, n):
    _say("lanczos n=%d" % n)
    rng = np.random.default_rng(100 + n)
    X = rng.uniform(-1, 1, (n, n))
    S = X @ X.T + X.T @ X
    start = rng.uniform(-1, 1, n)
    e_o, psi_o, it_o, conv_o = ora.lanczos_ev(S, start)
    e_g, psi_g, it_g = u10.lanczosEigh(S, start)
    assert conv_o
    scale = np.abs(np.linalg.eigvalsh(S)).max()
    assert abs(e_g - e_o) <= 1e-12 * scale
    assert abs(it_g - it_o) <= 3
    # same Ritz vector and sign convention (<start|psi> >= 0); the vector is
    # as accurate as the stopping rule makes it, so compare at that level
    assert np.dot(psi_g, start) >= 0
    rg = np.linalg.norm(S @ psi_g - e_g * psi_g) / np.linalg.norm(psi_g)
    assert rg <= 1e-4 * scale
    cos = np.dot(psi_g, psi_o) / (np.linalg.norm(psi_g) * np.linalg.norm(psi_o))
    assert cos > 1 - 1e-8


def test_lanczos_four_site_ring(u10, ora):
    H = ora.heisenberg_ring(4)
    e, psi, it = u10.lanczosEigh(H, np.random.default_rng(2).uniform(-1, 1, 16))
    assert abs(e + 2.0) < 1e-13
    assert np.linalg.norm(H @ psi + 2.0 * psi) < 1e-12


def test_lanczos_errors_and_edge_cases(u10):
    with pytest.raises(RuntimeError, match="greater than 2"):
        u10.lanczosEigh(np.eye(4), np.ones(4), max_iter=2)
    with pytest.raises(RuntimeError, match="non-square"):
        u10.lanczosEigh(np.ones((3, 4)), np.ones(3))
    with pytest.raises(ValueError):
        u10.lanczosEigh(np.eye(4), np.ones(3))
    # start vector = eigenvector: the reference's E0 = 0, psi unchanged
    e, psi, it = u10.lanczosEigh(np.diag([1.0, 2.0, 3.0]), [0.0, 1.0, 0.0])
    assert (e, it) == (0.0, 1)
    assert np.array_equal(psi, [0.0, 1.0, 0.0])
    rng = np.random.default_rng(7)
    X = rng.uniform(-1, 1, (200, 200))
    with pytest.raises(RuntimeError, match="fails in converging"):
        u10.lanczosEigh(X + X.T, rng.uniform(-1, 1, 200), max_iter=3)


def test_network_print_matches_network_h_known_answer(u10, ora, tmp_path):
    """The product's greedy planner and its operator<< reproduce the tree,
    node element counts and label order printed in Network.h:145-193."""
    from netcases import NETWORK_H_NET, NETWORK_H_PRINT, network_h_bonds
    p = tmp_path / "network_h.net"
    p.write_text(NETWORK_H_NET)
    names = [l.split(":")[0] for l in NETWORK_H_NET.splitlines()][:8]
    net = u10.Network(str(p))
    spec = ora.Network(str(p))
    for nm, labs, r in zip(spec.names[:-1], spec.label_arr[:-1], spec.Rnums[:-1]):
        T = ora.UniTensor(network_h_bonds(ora, labs, r), nm)
        net.putTensor(nm, to_product(u10, T))
    assert names == spec.names[:-1]
    assert str(net) == NETWORK_H_PRINT


def test_batch_tebd_reports_a_nonconverged_svd(u10, tn, golden):
    """BatchTebd.wait checks the SVD plans of both updates over every step
    launched: with the sweep cap forced to one sweep from the host, the first
    (cold) update cannot converge and the step raises instead of returning."""
    _say("batch tebd with sweep cap 1")
    tn.call("tn_svd_set_max_sweeps", 1)
    try:
        B = u10.BatchTebd([16, 24], [1.0, 0.5], [3, 4], 0.01, golden)
        with pytest.raises(u10.Uni10Error, match="did not converge"):
            B.step(1)
    finally:
        tn.call("tn_svd_set_max_sweeps", 0)
    B2 = u10.BatchTebd([16, 24], [1.0, 0.5], [3, 4], 0.01, golden)
    B2.step(2)


def test_otimes_and_operator_mul_values(u10, ora):
    """otimes and operator*(UniTensor, UniTensor) against the oracle: the
    tensor product is bit-exact (one product per element), operator* is
    contract of copies within 1e-12 and leaves its operands untouched."""
    A, B = _pair_u1(ora, 3)
    PA, PB = to_product(u10, A), to_product(u10, B)
    O = ora.otimes(A, B)
    PO = u10.otimes(PA, PB)
    assert PO.inBondNum() == O.RBondNum and PO.label() == O.labels
    assert PO.getElem().tobytes() == O.elem.tobytes()
    C = ora.contract(A.copy(), B.copy(), True)
    PC = PA * PB
    assert PC.label() == C.labels
    assert np.linalg.norm(PC.getElem() - C.elem) <= 1e-12 * np.linalg.norm(C.elem)
    assert PA.getElem().tobytes() == A.elem.tobytes() and PA.label() == A.labels


@pytest.mark.parametrize("case", [((5, 7), False, (7, 4), False), ((6, 6), True, (6, 3), False),
                                  ((4, 6), True, (6, 5), False), ((5, 6), False, (6, 6), True),
                                  ((5, 6), False, (6, 4), True), ((6, 6), True, (6, 6), True),
                                  ((3, 5), True, (5, 4), True)])
def test_rdotr_every_branch_matches_oracle(u10, ora, case):
    """RDotR: matrixMul, diagRowMul, diagColMul and the diagonal product, on
    square and rectangular diagonal operands (BlockTools.cpp:38-76)."""
    sa, da, sb, db = case
    _say("RDotR %s" % (case,))
    rng = np.random.default_rng(sum(sa) + 10 * sum(sb))
    a = rng.uniform(-1, 1, min(sa) if da else sa[0] * sa[1])
    b = rng.uniform(-1, 1, min(sb) if db else sb[0] * sb[1])
    c, shape, diag = u10.rdotr(a, sa, da, b, sb, db)
    w, wshape, wdiag = ora.rdotr(a, sa, da, b, sb, db)
    assert shape == wshape and diag == wdiag
    if da or db:
        assert c.tobytes() == np.ascontiguousarray(w).tobytes()    # one product per element
    else:
        assert np.linalg.norm(c - w) <= 1e-13 * np.linalg.norm(w)
    with pytest.raises(u10.Uni10Error, match="do not match"):
        u10.rdotr(a, sa, da, b, (sb[0] + 1, sb[1]), db)


def test_network_file_errors_match_oracle(u10, ora, tmp_path):
    """The .net errors the oracle raises (missing TOUT, imbalanced ORDER
    brackets, an unterminated last line dropped, unknown or repeated names
    in ORDER) are errors of the product's parser too."""
    cases = ["A: 1; 2\nB: 2; 3\n",
             "A: 1; 2\nB: 2; 3\nTOUT: 1; 3\nORDER: ((A B)\n",
             "A: 1; 2\nB: 2; 3\nTOUT: 1; 3",
             "A: 1; 2\nB: 2; 3\nTOUT: 1; 3\nORDER: A C\n",
             "A: 1; 2\nB: 2; 3\nTOUT: 1; 3\nORDER: A A\n",
             "A: 1; 2\nTOUT: 1; 2\n"]
    for k, text in enumerate(cases):
        p = tmp_path / ("bad%d.net" % k)
        p.write_text(text)
        with pytest.raises(ValueError):
            ora.Network(str(p))
        with pytest.raises(u10.Uni10Error):
            u10.Network(str(p))
    good = tmp_path / "good.net"
    good.write_text("A: 1; 2\nB: 2; 3\nTOUT: 1; 3\nORDER: B A\n")
    u10.Network(str(good))
    ora.Network(str(good))


def test_contract_dimension_mismatch_is_an_error(u10, ora):
    """Contracting bonds of different dimensions raises in both (the
    reference's "Cannot contract two bonds having different dimensions")."""
    A = u10.UniTensor([u10.Bond(u10.BD_IN, 4), u10.Bond(u10.BD_OUT, 5)])
    B = u10.UniTensor([u10.Bond(u10.BD_IN, 6), u10.Bond(u10.BD_OUT, 3)])
    A.setElem(np.ones(20))
    B.setElem(np.ones(18))
    A.setLabel([1, 2])
    B.setLabel([2, 3])
    with pytest.raises(u10.Uni10Error, match="different dimensions"):
        u10.contract(A, B)
    oA = ora.UniTensor([ora.Bond(ora.BD_IN, 4), ora.Bond(ora.BD_OUT, 5)])
    oB = ora.UniTensor([ora.Bond(ora.BD_IN, 6), ora.Bond(ora.BD_OUT, 3)])
    oA.set_elem(np.ones(20))
    oB.set_elem(np.ones(18))
    oA.set_label([1, 2])
    oB.set_label([2, 3])
    with pytest.raises(ValueError):
        ora.contract(oA, oB)


def test_batch_tebd_512_chains_configured_size_matches_oracle(u10, ora, golden):
    """The configured batch: 512 spin-1/2 XXZ chains, Delta ~ U[0, 2],
    chi cycling over 16, 32, ..., 256, chains 0-255 U(1) and 256-511 dense,
    stepped twice with graph replay as in the benchmark.  A stratified
    sample of 16 chains (every chi band; U(1) chains with even bands, dense
    chains with odd ones) follows the oracle's single-chain iTEBD (dense) or
    per-charge-block update (U(1)) to 1e-10 in energy, both Schmidt spectra
    and the discarded weight."""
    n, tau = 512, 0.01
    chis = list(range(16, 257, 16))
    chi = [chis[i % 16] for i in range(n)]
    sym = [1 if i < n // 2 else 0 for i in range(n)]
    delta = np.random.default_rng(4).uniform(0.0, 2.0, n)
    seeds = [1000 + 4 * i for i in range(n)]
    _say("512-chain batch: build")
    B = u10.BatchTebd(chi, delta, seeds, tau, golden, sym=sym)
    B.set_graph(True)
    B.step(1)
    B.step(1)
    _say("512-chain batch: 2 steps done")
    sample = [2 * k for k in range(8)] + [n // 2 + 2 * k + 1 for k in range(8)]
    assert sorted(chi[i] for i in sample) == chis
    for i in sample:
        if sym[i]:
            O = ora.ITEBDU1(chi[i], 2, float(delta[i]), tau, seeds[i], golden)
        else:
            O = ora.ITEBD(chi[i], 2, ora.xxz_two_site_d(float(delta[i]), 2), tau, seeds[i], golden)
        for _ in range(2):
            O.step()
        la, lb = (O.lam[0], O.lam[1]) if sym[i] else (O.La, O.Lb)
        e, eo = B.energy(i), O.energy()
        _say("chain %d chi=%d %s: %.15f oracle %.15f" % (i, chi[i], "U(1)" if sym[i] else "dense", e, eo))
        assert abs(e - eo) <= 1e-10
        assert np.max(np.abs(B.lam(i, 0) - la)) <= 1e-10
        assert np.max(np.abs(B.lam(i, 1) - lb)) <= 1e-10
        assert abs(B.trunc_err(i) - O.trunc_err) <= 1e-10


def test_network_graph_replay_matches_eager(u10, ora, golden):
    """Network.setGraph: launches replay a captured graph.  On fresh inputs
    put before every launch (with other device work queued in between), on a
    change of bond dimension (a new capture) and back, the replayed result
    equals the eager launch's bit for bit."""
    E = u10.Network(os.path.join(golden, "itebd.net"))
    G = u10.Network(os.path.join(golden, "itebd.net"))
    G.setGraph(True)
    d = 2
    U = u10.UniTensor([u10.Bond(u10.BD_IN, d)] * 2 + [u10.Bond(u10.BD_OUT, d)] * 2)
    U.putBlock(ora.exph(-0.01, ora.heisenberg_two_site(0.8)))
    for it, chi in enumerate((12, 12, 12, 20, 12)):
        b3 = [u10.Bond(u10.BD_IN, chi), u10.Bond(u10.BD_IN, chi), u10.Bond(u10.BD_OUT, d)]
        A, B = u10.UniTensor(b3), u10.UniTensor(b3)
        A.setElem(ora.splitmix64_uniform(100 + it, A.elemNum()))
        B.setElem(ora.splitmix64_uniform(200 + it, B.elemNum()))
        for net in (E, G):
            net.putTensor("ALa", A)
            net.putTensor("BLb", B)
            net.putTensor("expH", U)
        re_ = E.launch()
        junk = u10.contract(A.copy().setLabel([1, 2, 3]), B.copy().setLabel([5, 4, 3]))   # other work between
        rg = G.launch()
        _say("graph replay iteration %d chi=%d" % (it, chi))
        assert rg.label() == re_.label() and rg.inBondNum() == re_.inBondNum()
        assert rg.getElem().tobytes() == re_.getElem().tobytes()
        del junk


def test_network_graph_replay_u1_network(u10, ora, tmp_path):
    """The Network.h network with U(1) bonds (a scalar result) in graph mode:
    two replays on different inputs equal the eager launches."""
    from netcases import NETWORK_H_NET, network_h_bonds
    p = tmp_path / "network_h.net"
    p.write_text(NETWORK_H_NET)
    spec = ora.Network(str(p))
    E, G = u10.Network(str(p)), u10.Network(str(p))
    G.setGraph(True)
    for rep in range(2):
        rng = np.random.default_rng(60 + rep)
        for nm, labs, r in zip(spec.names[:-1], spec.label_arr[:-1], spec.Rnums[:-1]):
            T = ora.UniTensor(network_h_bonds(ora, labs, r), nm)
            T.set_elem(rng.uniform(-1, 1, T.elem_num()))
            P = to_product(u10, T)
            E.putTensor(nm, P)
            G.putTensor(nm, P)
        a, b = E.launch(), G.launch()
        assert a.getElem().tobytes() == b.getElem().tobytes()


def test_batch_tebd_exact_bf16_mode_is_fp64_grade(u10, golden):
    """Batched iTEBD with the exact split-BF16 sector GEMMs against the same
    batch in FP64: energies per site within 1e-10 after 200 steps (the
    FP32 and split-BF16 forms with FP32 sums differ by ~1e-7 here)."""
    chi, delta, seed = [8, 16, 24, 32], [0.5, 1.0, 1.5, 1.0], [21, 22, 23, 24]
    e = {}
    for mode in (0, 4):
        _say("batched TEBD mode %d" % mode)
        B = u10.BatchTebd(chi, delta, seed, 0.01, golden, mode)
        B.set_graph(True)
        B.step(200)
        e[mode] = np.array([B.energy(i) for i in range(len(chi))])
    _say("exact vs f64 energies: %s" % np.abs(e[4] - e[0]))
    assert np.max(np.abs(e[4] - e[0])) <= 1e-10


def test_batch_tebd_inactive_chains_are_skipped(u10, golden):
    """BatchTebd.set_active: inactive chains are skipped on the device
    (one-workgroup and block-Jacobi SVDs, GEMMs and vector ops), their state
    stays as it was, the active chains are unaffected bit for bit, and a
    chain reactivated later continues exactly as if it had never paused."""
    chi, delta, seed = [8, 16, 24, 72, 12, 80], [1.0, 0.6, 1.4, 0.9, 0.3, 1.2], [31, 32, 33, 34, 35, 36]
    sym = [0, 1, 0, 0, 1, 1]
    def eng():
        B = u10.BatchTebd(chi, delta, seed, 0.01, golden, sym=sym)
        B.set_graph(True)
        return B
    A, B, R = eng(), eng(), eng()
    for E in (A, B, R):
        E.step(2)
    off = (1, 3, 5)
    for i in off:
        B.set_active(i, False)
    held = {i: (B.lam(i, 0).copy(), B.lam(i, 1).copy()) for i in off}
    A.step(3)
    B.step(3)
    _say("masked steps done")
    for i in range(len(chi)):
        if i in off:
            assert B.lam(i, 0).tobytes() == held[i][0].tobytes() and B.lam(i, 1).tobytes() == held[i][1].tobytes()
        else:
            assert B.lam(i, 0).tobytes() == A.lam(i, 0).tobytes()
            assert B.lam(i, 1).tobytes() == A.lam(i, 1).tobytes()
    for i in off:
        B.set_active(i, True)
    B.step(3)
    R.step(3)
    for i in off:   # 2 + 3 active steps each
        assert B.lam(i, 0).tobytes() == R.lam(i, 0).tobytes()
        assert B.lam(i, 1).tobytes() == R.lam(i, 1).tobytes()
        assert B.energy(i) == R.energy(i)


# ---------------------------------------------------------------------------
# Complex tensors (UniTensorComplex.cpp): checked against the real path,
# which the oracle pins: a permute moves real and imaginary parts as the
# real permute moves each, a contraction is the four real contractions of
# the parts (UniTensorTools.cpp:58-82 for the real-complex promotion).
def _twin(ora, T, seed):
    rng = np.random.default_rng(seed)
    U = ora.UniTensor([b.copy() for b in T.bonds])
    U.set_raw_elem(rng.uniform(-1, 1, [b.dim for b in T.bonds]))
    U.set_label(list(T.labels))
    return U


def to_product_c(u10, Tr, Ti):
    bonds = [u10.Bond(b.type, [q.triple() for q in b.Qlist()]) for b in Tr.bonds]
    P = u10.UniTensor(bonds, Tr.name, dtype=complex)
    P.setLabel(Tr.labels)
    P.setElem(Tr.elem + 1j * Ti.elem)
    return P


def same_layout_c(P, T):
    pb = [(q.triple(), r, c // 2, o // 2) for q, r, c, o in P.blocks()]
    tb = [(q.triple(), r, c, o) for q, (r, c, o) in T.layout.blocks.items()]
    return pb == tb


@pytest.mark.parametrize("seed", range(8))
def test_complex_permute_moves_both_parts_bitexact(u10, ora, seed):
    rng = np.random.default_rng(900 + seed)
    Tr = _random_u1(ora, 40 + seed, n_in=int(rng.integers(1, 3)), n_out=int(rng.integers(1, 3)))
    Ti = _twin(ora, Tr, 60 + seed)
    P = to_product_c(u10, Tr, Ti)
    assert P.is_complex and P.elemNum() == Tr.elem.size and same_layout_c(P, Tr)
    nb = len(Tr.bonds)
    perm = [int(v) for v in rng.permutation(nb)]
    rb = int(rng.integers(0, nb + 1))
    labels = [Tr.labels[p] for p in perm]
    _say("complex permute seed=%d perm=%s rb=%d" % (seed, perm, rb))
    try:
        Ro = Tr.copy().permute(labels, rb)
        Io = Ti.copy().permute(labels, rb)
    except ValueError:
        with pytest.raises(u10.Uni10Error):
            P.permute(labels, rb)
        return
    P.permute(labels, rb)
    assert P.label() == labels and P.inBondNum() == rb and same_layout_c(P, Ro)
    e = P.getElem()
    assert e.real.tobytes() == Ro.elem.tobytes() and e.imag.tobytes() == Io.elem.tobytes()
    raw = P.getRawElem()
    assert np.array_equal(raw.real.ravel(), Ro.get_raw_elem().ravel())
    assert np.array_equal(raw.imag.ravel(), Io.get_raw_elem().ravel())


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("fast", [False, True])
def test_complex_contract_is_the_real_composition(u10, ora, seed, fast):
    A, B = _pair_u1(ora, 70 + seed)
    Ai, Bi = _twin(ora, A, 80 + seed), _twin(ora, B, 90 + seed)
    PA, PB = to_product_c(u10, A, Ai), to_product_c(u10, B, Bi)
    rr = ora.contract(A.copy(), B.copy())
    ii = ora.contract(Ai.copy(), Bi.copy())
    ri = ora.contract(A.copy(), Bi.copy())
    ir = ora.contract(Ai.copy(), B.copy())
    ref = (rr.elem - ii.elem) + 1j * (ri.elem + ir.elem)
    C = ora.contract(A, B, fast)          # labels, layout and the inputs' state
    PC = u10.contract(PA, PB, fast)
    assert PC.is_complex and PC.label() == C.labels and same_layout_c(PC, C)
    err = np.linalg.norm(PC.getElem() - ref) / np.linalg.norm(ref)
    _say("complex contract seed=%d fast=%s err=%.2e" % (seed, fast, err))
    assert err <= 1e-12
    assert PA.label() == A.labels and PB.label() == B.labels
    assert PA.getElem().real.tobytes() == A.elem.tobytes()


def test_complex_contract_promotes_a_real_operand(u10, ora):
    """real x complex: the real operand is promoted on a copy (RtoC) and
    stays as it was."""
    A, B = _pair_u1(ora, 101)
    Bi = _twin(ora, B, 102)
    PA, PB = to_product(u10, A), to_product_c(u10, B, Bi)
    before = PA.getElem().tobytes()
    ref = ora.contract(A.copy(), B.copy()).elem + 1j * ora.contract(A.copy(), Bi.copy()).elem
    PC = u10.contract(PA, PB)
    assert PC.is_complex and not PA.is_complex and PA.getElem().tobytes() == before
    assert np.linalg.norm(PC.getElem() - ref) <= 1e-12 * np.linalg.norm(ref)
    PC2 = u10.contract(PB, PA)            # complex x real
    ref2 = ora.contract(B.copy(), A.copy()).elem + 1j * ora.contract(Bi.copy(), A.copy()).elem
    assert np.linalg.norm(PC2.getElem() - ref2) <= 1e-12 * np.linalg.norm(ref2)


def _dense_pair(u10, a, labels, dtype=float):
    T = u10.UniTensor([u10.Bond(u10.BD_IN, a.shape[0]), u10.Bond(u10.BD_OUT, a.shape[1])], dtype=dtype)
    T.setRawElem(a)
    T.setLabel(labels)
    return T


def test_complex_outer_product_and_dense_contract(u10):
    """Dense bonds: against numpy's complex matrix product; no shared
    bonds: the real path's bond order, values the real composition."""
    rng = np.random.default_rng(7)
    a = rng.uniform(-1, 1, (3, 4)) + 1j * rng.uniform(-1, 1, (3, 4))
    b = rng.uniform(-1, 1, (4, 5)) + 1j * rng.uniform(-1, 1, (4, 5))
    C = u10.contract(_dense_pair(u10, a, [1, 2], complex), _dense_pair(u10, b, [2, 3], complex))
    assert C.label() == [1, 3]
    assert np.max(np.abs(C.getRawElem() - a @ b)) <= 1e-13
    O = u10.contract(_dense_pair(u10, a, [1, 2], complex), _dense_pair(u10, b, [5, 6], complex))

    def real_outer(x, y):
        return u10.contract(_dense_pair(u10, x, [1, 2]), _dense_pair(u10, y, [5, 6]))

    rr, ii = real_outer(a.real, b.real), real_outer(a.imag, b.imag)
    ri, ir = real_outer(a.real, b.imag), real_outer(a.imag, b.real)
    assert O.label() == rr.label() and O.inBondNum() == rr.inBondNum()
    ref = (rr.getRawElem() - ii.getRawElem()) + 1j * (ri.getRawElem() + ir.getRawElem())
    assert np.max(np.abs(O.getRawElem() - ref)) <= 1e-13


def test_complex_blocks_and_elements(u10, ora):
    Tr = _random_u1(ora, 111)
    Ti = _twin(ora, Tr, 112)
    P = to_product_c(u10, Tr, Ti)
    for q, (r, c, o) in Tr.layout.blocks.items():
        blk = P.getBlock(u10.Qnum(*q.triple()))
        exp = (Tr.elem[o:o + r * c] + 1j * Ti.elem[o:o + r * c]).reshape(r, c)
        assert np.array_equal(blk, exp)
        P.putBlock(2.0 * blk, u10.Qnum(*q.triple()))
        assert np.array_equal(P.getBlock(u10.Qnum(*q.triple())), 2.0 * exp)
    assert abs(P.norm() - 2.0 * np.sqrt(np.sum(Tr.elem ** 2) + np.sum(Ti.elem ** 2))) <= 1e-12 * P.norm()


def test_complex_svd_eigh_through_matrix(u10):
    rng = np.random.default_rng(5)
    a = rng.uniform(-1, 1, (7, 5)) + 1j * rng.uniform(-1, 1, (7, 5))
    U, S, VT = u10.svd(a)
    assert np.max(np.abs(S - np.linalg.svd(a, compute_uv=False))) <= 1e-13 * S[0]
    assert np.linalg.norm(U @ np.diag(S) @ VT - a) <= 1e-13 * np.linalg.norm(a)
    h = a.conj().T @ a + np.diag(rng.uniform(-3, 3, 5))
    w, E = u10.eigh(h)
    assert np.max(np.abs(w - np.linalg.eigvalsh(h))) <= 1e-12 * np.abs(w).max()
    assert np.linalg.norm(E.conj().T @ np.diag(w) @ E - h) <= 1e-11 * np.linalg.norm(h)


def test_complex_tensor_unsupported_paths_fail(u10, golden, tmp_path):
    P = u10.UniTensor([u10.Bond(u10.BD_IN, 2), u10.Bond(u10.BD_OUT, 2)], dtype=complex)
    P.setRawElem(np.arange(4) + 1j)
    with pytest.raises(u10.Uni10Error):
        P.at([0, 0])
    with pytest.raises(u10.Uni10Error):
        P.save(str(tmp_path / "c.ut"))
    net = u10.Network(os.path.join(golden, "Simple.net"))
    with pytest.raises(u10.Uni10Error):
        net.putTensor("A", P)
