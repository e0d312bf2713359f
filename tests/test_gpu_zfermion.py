"""GPU parity of the fermionic paths (swap-gate signs) against the oracle.

Bars: fermionic permutes and addGate are bit-exact against the oracle (the
signs only flip copies); a fermionic network contraction is within 1e-12 of
the oracle's and does not depend on the contraction tree.

Creating a fermion-odd Qnum switches the process to fermionic bookkeeping
(Qnum.cpp:33,54-55, kept by the product), so this file runs after the other
GPU parity files; with all-even charges every sign is +1 and the other
results are unchanged.
"""
import numpy as np
import pytest

from netcases import CHAIN3_NET, FERMI_ORDERS, NETWORK_H_NET, fermi_tensor, network_h_fermionic_bonds

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def u10(tn):
    from trim_amd import uni10
    uni10.load()
    uni10.set_gemm_mode(uni10.GEMM_F64)
    return uni10


@pytest.fixture
def ferm(ora, monkeypatch):
    monkeypatch.setattr(ora.Qnum, "fermionic", True)
    return ora


def to_product(u10, T):
    bonds = [u10.Bond(b.type, [q.triple() for q in b.Qlist()]) for b in T.bonds]
    P = u10.UniTensor(bonds, T.name)
    P.setLabel(T.labels)
    P.setElem(T.elem)
    return P


@pytest.mark.parametrize("seed", range(24))
def test_fermionic_permute_bitexact(u10, ferm, seed):
    ora = ferm
    rng = np.random.default_rng(900 + seed)
    nb = int(rng.integers(2, 6))
    T = fermi_tensor(ora, rng, nb, int(rng.integers(0, nb + 1)))
    if T is None:
        pytest.skip("no charge-conserving layout drawn")
    l1 = [T.labels[p] for p in rng.permutation(nb)]
    l2 = [T.labels[p] for p in rng.permutation(nb)]
    r1, r2 = int(rng.integers(0, nb + 1)), int(rng.integers(0, nb + 1))
    _say("fermionic permute seed=%d nb=%d rb=%d -> %s/%d -> %s/%d" % (seed, nb, T.RBondNum, l1, r1, l2, r2))
    try:
        O1 = T.copy().permute(l1, r1)
        O2 = O1.copy().permute(l2, r2)
    except ValueError:
        pytest.skip("an intermediate layout carries no common charge")
    P = to_product(u10, T)
    P.permute(l1, r1)
    assert P.getElem().tobytes() == O1.elem.tobytes()
    P.permute(l2, r2)
    assert P.getElem().tobytes() == O2.elem.tobytes()
    P.permute(T.labels, T.RBondNum)
    assert P.getElem().tobytes() == T.elem.tobytes()


@pytest.mark.parametrize("seed", range(8))
def test_add_gate_and_ex_swap_bitexact(u10, ferm, seed):
    ora = ferm
    rng = np.random.default_rng(950 + seed)
    nb = int(rng.integers(2, 5))
    T = fermi_tensor(ora, rng, nb, int(rng.integers(0, nb + 1)))
    if T is None:
        pytest.skip("no charge-conserving layout drawn")
    gates = [tuple(int(v) for v in rng.choice(nb, 2, replace=False)) for _ in range(int(rng.integers(1, 4)))]
    _say("addGate seed=%d nb=%d gates=%s" % (seed, nb, gates))
    P = to_product(u10, T)
    P.addGate(gates)
    T.add_gate(gates)
    assert P.getElem().tobytes() == T.elem.tobytes()
    # exSwap against a second tensor sharing two labels
    S = T.copy()
    S.set_label([T.labels[0], 99] + list(range(200, 200 + nb - 2)))
    Q = to_product(u10, S)
    assert P.exSwap(Q) == T.ex_swap(S)


def _product_network(u10, ora, path, tensors):
    net = u10.Network(path)
    for nm, T in tensors.items():
        net.putTensor(nm, to_product(u10, T))
    return net


def test_fermionic_chain_network_matches_oracle(u10, ferm, tmp_path):
    ora = ferm
    Q = ora.Qnum
    qs = [Q(0, ora.PRT_EVEN, ora.PRTF_EVEN), Q(0, ora.PRT_EVEN, ora.PRTF_ODD), Q(0, ora.PRT_EVEN, ora.PRTF_ODD)]
    rng = np.random.default_rng(5)
    ts = {}
    for nm in "ABC":
        T = ora.UniTensor([ora.Bond(ora.BD_IN, qs), ora.Bond(ora.BD_OUT, qs)], nm)
        T.set_elem(rng.uniform(-1, 1, T.elem_num()))
        ts[nm] = T
    want = ts["A"].get_raw_elem() @ ts["B"].get_raw_elem() @ ts["C"].get_raw_elem()
    for order in ("((A C) B)", "((A B) C)"):
        p = tmp_path / "chain3.net"
        p.write_text(CHAIN3_NET % order)
        _say("fermionic chain network order %s" % order)
        net = _product_network(u10, ora, str(p), ts)
        R = net.launch()
        assert np.allclose(R.getRawElem(), want, rtol=0, atol=1e-14)
        # a relaunch applies no gate twice (that would negate the odd block);
        # its products differ in shape (contract(fast=true) leaves the leaf
        # tensors permuted, UniTensorTools.cpp:178-179), so compare values
        assert np.allclose(net.launch().getRawElem(), want, rtol=0, atol=1e-14)


def test_fermionic_network_h_matches_oracle_under_every_order(u10, ferm, tmp_path):
    """The Network.h network with fermion-parity charges under three trees
    (different swap gates): the product equals the oracle to 1e-12 and all
    trees give the same scalar."""
    ora = ferm
    vals = []
    for k, order in enumerate(FERMI_ORDERS):
        p = tmp_path / ("f%d.net" % k)
        p.write_text(NETWORK_H_NET.replace("ORDER: " + FERMI_ORDERS[0], "ORDER: " + order))
        spec = ora.Network(str(p))
        rng = np.random.default_rng(1)
        ts = {}
        for nm, labs, r in zip(spec.names[:-1], spec.label_arr[:-1], spec.Rnums[:-1]):
            T = ora.UniTensor(network_h_fermionic_bonds(ora, labs, r), nm)
            T.set_elem(rng.uniform(-1, 1, T.elem_num()))
            ts[nm] = T
            spec.put_tensor(nm, T)
        want = float(spec.launch().elem[0])
        _say("fermionic Network.h order %s" % order)
        got = float(_product_network(u10, ora, str(p), ts).launch().getElem()[0])
        assert abs(got - want) <= 1e-12 * max(1.0, abs(want))
        vals.append(got)
    assert max(vals) - min(vals) <= 1e-12 * max(1.0, abs(vals[0]))


def test_fermionic_network_graph_replay_matches_eager(u10, ferm, tmp_path):
    """Graph replay of a fermionic network applies the swap gates inside the
    captured launch: equal to the eager launch bit for bit, on two inputs."""
    ora = ferm
    Q = ora.Qnum
    qs = [Q(0, ora.PRT_EVEN, ora.PRTF_EVEN), Q(0, ora.PRT_EVEN, ora.PRTF_ODD), Q(0, ora.PRT_EVEN, ora.PRTF_ODD)]
    p = tmp_path / "chain3.net"
    p.write_text(CHAIN3_NET % "((A C) B)")
    E, G = u10.Network(str(p)), u10.Network(str(p))
    G.setGraph(True)
    for rep in range(2):
        rng = np.random.default_rng(70 + rep)
        for nm in "ABC":
            T = ora.UniTensor([ora.Bond(ora.BD_IN, qs), ora.Bond(ora.BD_OUT, qs)], nm)
            T.set_elem(rng.uniform(-1, 1, T.elem_num()))
            P = to_product(u10, T)
            E.putTensor(nm, P)
            G.putTensor(nm, P)
        a, b = E.launch(), G.launch()
        assert a.getElem().tobytes() == b.getElem().tobytes()
