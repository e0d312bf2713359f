"""Network cases shared by the oracle tests and the GPU parity tests.

NETWORK_H_NET / NETWORK_H_PRINT: the network and the printed contraction
tree documented in the reference's Network.h:145-193, a known answer for the
greedy planner (tree shape, node element counts, label order).
"""
import numpy as np

NETWORK_H_NET = """W1: -1; 0 1 3
W2: -2; 7 10 11
U: 3 7; 4 8
Ob: 1 4; 2 5
UT: 5 8; 6 9
W1T: 0 2 6; -3
W2T: 9 10 11; -4
Rho: -3 -4; -1 -2
TOUT:
ORDER: W1 W1T W2 W2T U Ob UT Rho
"""

# Network.h:163-172 (tensor list) and :181-195 (the tree with element counts)
NETWORK_H_PRINT = """
W1: i[-1] o[0, 1, 3]
W2: i[-2] o[7, 10, 11]
U: i[3, 7] o[4, 8]
Ob: i[1, 4] o[2, 5]
UT: i[5, 8] o[6, 9]
W1T: i[0, 2, 6] o[-3]
W2T: i[9, 10, 11] o[-4]
Rho: i[-3, -4] o[-1, -2]
TOUT: 

*(1): 
|   *(70): 7, 9, -4, -2, 
|   |   *(70): -1, 7, 9, -3, 
|   |   |   *(70): -1, 0, 7, 2, 6, 9, 
|   |   |   |   W1(20): -1, 0, 1, 3, 
|   |   |   |   *(20): 3, 7, 1, 2, 6, 9, 
|   |   |   |   |   *(20): 3, 7, 8, 1, 2, 5, 
|   |   |   |   |   |   U(6): 3, 7, 4, 8, 
|   |   |   |   |   |   Ob(6): 1, 4, 2, 5, 
|   |   |   |   |   UT(6): 5, 8, 6, 9, 
|   |   |   W1T(20): 0, 2, 6, -3, 
|   |   Rho(924): -3, -4, -1, -2, 
|   *(70): -2, 7, 9, -4, 
|   |   W2(20): -2, 7, 10, 11, 
|   |   W2T(20): 9, 10, 11, -4, 
"""


def network_h_bonds(mod, labels, rnum):
    """Bonds that reproduce the element counts Network.h prints: internal
    legs are spin 1/2 (Sz charges +1, -1), external legs (negative labels)
    three fused spins 1/2 (charges 3, 1, -1, -3 with degeneracies 1, 3, 3, 1).
    Then W1 = 1 + 9 + 9 + 1 = 20, U = 1 + 4 + 1 = 6 and Rho = sum C(6,k)^2 = 924."""
    Q = mod.Qnum
    out = []
    for i, l in enumerate(labels):
        tp = mod.BD_IN if i < rnum else mod.BD_OUT
        if l < 0:
            qs = [Q(3)] + [Q(1)] * 3 + [Q(-1)] * 3 + [Q(-3)]
        else:
            qs = [Q(1), Q(-1)]
        out.append(mod.Bond(tp, qs))
    return out


def network_h_fermionic_bonds(mod, labels, rnum):
    """The Network.h bonds with fermion parity: a spin-1/2 leg carries
    (Sz +1, odd), (Sz -1, even); the three-spin external leg carries the
    parity of its number of up spins."""
    Q = mod.Qnum
    out = []
    for i, l in enumerate(labels):
        tp = mod.BD_IN if i < rnum else mod.BD_OUT
        if l < 0:
            qs = [Q(3, 0, 1)] + [Q(1, 0, 0)] * 3 + [Q(-1, 0, 1)] * 3 + [Q(-3, 0, 0)]
        else:
            qs = [Q(1, 0, 1), Q(-1, 0, 0)]
        out.append(mod.Bond(tp, qs))
    return out


FERMI_ORDERS = ("W1 W1T W2 W2T U Ob UT Rho", "((((((W1 U) Ob) UT) W1T) Rho) (W2 W2T))",
                "(((U Ob) (UT W1)) ((W1T Rho) (W2 W2T)))")


CHAIN3_NET = "A: 1; 2\nB: 2; 3\nC: 3; 4\nTOUT: 1; 4\nORDER: %s\n"


def fermi_tensor(ora, rng, nb, rb):
    """Random bonds with charges (U1 in {-1,0,1}, fermion parity) and a
    random charge-conserving tensor."""
    for _ in range(50):
        bonds = []
        for b in range(nb):
            n = int(rng.integers(1, 4))
            qs = []
            for _ in range(n):
                u, f = int(rng.integers(-1, 2)), int(rng.integers(0, 2))
                qs += [ora.Qnum(u, ora.PRT_EVEN, f)] * int(rng.integers(1, 3))
            bonds.append(ora.Bond(ora.BD_IN if b < rb else ora.BD_OUT, qs))
        try:
            T = ora.UniTensor(bonds)
        except ValueError:
            continue
        T.set_elem(rng.uniform(-1, 1, T.elem_num()))
        T.set_label(list(range(10, 10 + nb)))
        return T
    return None
