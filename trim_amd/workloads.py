"""Synthetic workloads of the contraction hot path, built through the API.

* splitmix_uniform  - SplitMix64 -> uniform[-1,1) (x_i = mix(seed + i*golden), i >= 1)
* binomial_sectors  - U(1) sector dimensions: chi * C(2 qmax, q + qmax) / 4^qmax,
                      floored with a minimum of 1, the remainder handed out by
                      largest fractional part (ties: lower charge first), then
                      taken back from the largest sectors if the minimum
                      overshot.
* u1_two_site       - the U(1) two-site contraction: two charge-conserving MPS
                      tensors A(chi_l IN | s OUT, chi_r OUT) and a two-site gate
                      exp(-tau H) from a random Sz-conserving symmetric H;
                      Theta = (A1 . A2) . G
"""
from __future__ import annotations

from math import comb, floor
from typing import List, Sequence, Tuple

import numpy as np


def splitmix_uniform(seed: int, n: int) -> np.ndarray:
    M = np.uint64
    i = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = M(seed & ((1 << 64) - 1)) + i * M(0x9E3779B97F4A7C15)
        z = (z ^ (z >> M(30))) * M(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> M(27))) * M(0x94D049BB133111EB)
        z = z ^ (z >> M(31))
    return 2.0 * ((z >> M(11)).astype(np.float64) * (2.0 ** -53)) - 1.0


def binomial_sectors(chi: int, qmax: int) -> List[int]:
    w = [comb(2 * qmax, k) for k in range(2 * qmax + 1)]
    tot = float(4 ** qmax)
    raw = [chi * x / tot for x in w]
    dims = [max(1, floor(r)) for r in raw]
    deficit = chi - sum(dims)
    order = sorted(range(len(raw)), key=lambda i: (-(raw[i] - floor(raw[i])), i))
    k = 0
    while deficit > 0:
        dims[order[k % len(order)]] += 1
        deficit -= 1
        k += 1
    while deficit < 0:
        j = max(range(len(dims)), key=lambda i: dims[i])
        dims[j] -= 1
        deficit += 1
    return dims


def u1_bond(u10, tp: int, charges: Sequence[int], dims: Sequence[int]):
    qs = []
    for c, d in zip(charges, dims):
        qs += [u10.Qnum(c)] * int(d)
    return u10.Bond(tp, qs)


def sz_hamiltonian(seed: int) -> np.ndarray:
    """A random real symmetric two-site H that conserves total Sz (basis
    |uu>,|ud>,|du>,|dd>: blocks {uu}, {ud,du}, {dd}); entries from SplitMix64
    uniform[-1,1)."""
    r = splitmix_uniform(seed, 5)
    H = np.zeros((4, 4))
    H[0, 0], H[3, 3] = r[0], r[1]
    H[1, 1], H[2, 2] = r[2], r[3]
    H[1, 2] = H[2, 1] = r[4]
    return H


def u1_two_site(u10, chi: int = 1024, qmax: int = 8, seed: int = 11, tau: float = 0.01):
    """(A1, A2, G) as UniTensors, labelled for Theta = (A1 . A2) . G."""
    charges = list(range(-qmax, qmax + 1))
    dims = binomial_sectors(chi, qmax)
    chi_in = u1_bond(u10, u10.BD_IN, charges, dims)
    chi_out = u1_bond(u10, u10.BD_OUT, charges, dims)
    phys_out = u1_bond(u10, u10.BD_OUT, [1, -1], [1, 1])
    A1 = u10.UniTensor([chi_in, phys_out, chi_out], "A1")
    A2 = u10.UniTensor([chi_in, phys_out, chi_out], "A2")
    A1.setElem(splitmix_uniform(seed, A1.elemNum()))
    A2.setElem(splitmix_uniform(seed + 1, A2.elemNum()))
    A1.setLabel([-1, 10, 0])
    A2.setLabel([0, 11, -2])
    H = sz_hamiltonian(seed + 2)
    g = u10.exph(-tau, H)      # device eigh + exponential
    phys_in = u1_bond(u10, u10.BD_IN, [1, -1], [1, 1])
    G = u10.UniTensor([phys_in, phys_in, phys_out, phys_out], "G")
    G.setRawElem(g)
    G.setLabel([10, 11, 20, 21])
    return A1, A2, G, dims
