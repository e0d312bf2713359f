"""Continuous-batching request loop over the batched iTEBD engine.

Requests arrive on a Poisson clock and are of two kinds:

* ``tebd``: run ``steps`` iTEBD steps of one spin-1/2 XXZ chain (bond
  dimension ``chi``, anisotropy ``delta``, seeded start) and return its
  Schmidt spectra;
* ``contract``: contract two seeded (chi, d, chi) tensors over one bond
  through the drop-in ``contract`` and return the result's norm.

A request also names its physical dimension d (2: spin 1/2, 3: spin 1) and
symmetry (dense, or U(1) with conserved Sz: block-sparse tensors whose
sectors follow the U(1) chains' sector rule).  The server keeps one batched
engine per class (bond dimension, d, symmetry), each with a fixed number of
slots whose tensors stay resident in HBM; the slot count of each class is
what its share of the HBM budget holds.  Every tick it admits the requests
that have arrived into free slots (first come, first served per class),
and every class steps at its own pace: a class with a busy slot and no step
in flight launches one batched, graph-replayed step on its own stream (so
the classes' work overlaps on the device), and a class whose step has
finished retires its finished requests and admits new ones without waiting
for slower classes.  Requests that arrive while others run join at their
class's next step: a chain's
computation never depends on which other slots are busy (slots are
restarted with identity warm-start bases), so each result is the same alone
and in the mix.
"""
from __future__ import annotations

import dataclasses
import math
import time
from typing import Dict, List, Optional

import numpy as np

from . import uni10


@dataclasses.dataclass
class Request:
    rid: int
    t_arrival: float          # seconds from the start of the stream
    kind: str                 # "tebd" or "contract"
    chi: int
    delta: float
    seed: int
    steps: int
    d: int = 2
    sym: int = 0              # 0 dense, 1 U(1)
    t_admit: float = -1.0
    t_done: float = -1.0
    result: Optional[np.ndarray] = None


def poisson_requests(n: int, rate: float, seed: int, chi_classes, steps_max: int = 50,
                     tebd_fraction: float = 0.5, d_choices=(2,), u1_fraction: float = 0.0) -> List[Request]:
    """n requests with exponential inter-arrival times (mean 1/rate), bond
    dimension log-uniform over the classes' range and rounded to the nearest
    class, 1..steps_max steps, anisotropy uniform in [0, 2], d uniform over
    d_choices, U(1) with probability u1_fraction."""
    rng = np.random.default_rng(seed)
    classes = np.asarray(sorted(chi_classes), dtype=np.float64)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    lo, hi = math.log(classes[0]), math.log(classes[-1])
    out = []
    for i in range(n):
        chi = float(np.exp(rng.uniform(lo, hi)))
        chi = int(classes[np.argmin(np.abs(np.log(classes) - math.log(chi)))])
        kind = "tebd" if rng.random() < tebd_fraction else "contract"
        r = Request(rid=i, t_arrival=float(t[i]), kind=kind, chi=chi,
                    delta=float(rng.uniform(0.0, 2.0)), seed=int(10_000 + i),
                    steps=int(rng.integers(1, steps_max + 1)))
        if len(d_choices) > 1 or u1_fraction > 0.0:
            r.d = int(d_choices[int(rng.integers(0, len(d_choices)))])
            r.sym = int(rng.random() < u1_fraction)
        else:
            r.d = int(d_choices[0])
        out.append(r)
    return out


def slot_bytes(chi: int, d: int = 2, sym: int = 0) -> int:
    """HBM bytes one chain of the batched engine keeps resident (a dense
    chain's; a U(1) chain of the same chi and d holds less)."""
    D = chi * d
    doubles = 2 * chi * chi * d + 2 * chi + d ** 4 + 3 * chi * chi * d * d + 3 * D * D + D + 4
    return 8 * doubles


def _cls(c):
    """A class key (chi, d, sym); a bare chi means (chi, 2, 0)."""
    return (int(c), 2, 0) if isinstance(c, (int, np.integer)) else tuple(int(x) for x in c)


class Server:
    def __init__(self, classes, hbm_budget: float, netdir: Optional[str] = None, gemm_mode: int = 0,
                 max_slots: int = 256, tau: float = 0.01):
        """classes: (chi, d, sym) tuples (or bare chi for (chi, 2, dense))."""
        self.netdir = netdir
        self.gemm_mode = gemm_mode
        self.classes = sorted(_cls(c) for c in classes)
        share = hbm_budget / len(self.classes)
        self.engines: Dict[tuple, uni10.BatchTebd] = {}
        self.free: Dict[tuple, List[int]] = {}
        self.busy: Dict[tuple, Dict[int, list]] = {}
        for c in self.classes:
            chi, d, sym = c
            k = int(max(1, min(max_slots, share // slot_bytes(chi, d, sym))))
            eng = uni10.BatchTebd([chi] * k, [1.0] * k, [1 + s for s in range(k)], tau, netdir, gemm_mode,
                                  d=[d] * k, sym=[sym] * k)
            eng.set_graph(True)
            self.engines[c] = eng
            self.free[c] = list(range(k))
            self.busy[c] = {}
        self.chain_steps = 0

    def warmup(self):
        """Run the first (from-scratch) step and the graph capture of every
        engine before requests arrive; slots are restarted on admission."""
        for eng in self.engines.values():
            eng.step(2)

    def slots(self, cls) -> int:
        return len(self.engines[_cls(cls)].chi)

    def _contract(self, r: Request) -> np.ndarray:
        uni10.set_gemm_mode(self.gemm_mode)
        try:
            c, d = r.chi, r.d
            if r.sym:
                # A(a, s; m) with q_m = q_a + s: bond 1 sectors in, bond 0 sectors out
                q1, n1 = uni10.u1_bond_sectors(c, d, 1)
                q0, n0 = uni10.u1_bond_sectors(c, d, 0)
                qa = [uni10.Qnum(int(q)) for q, n in zip(q1, n1) for _ in range(int(n))]
                qm = [uni10.Qnum(int(q)) for q, n in zip(q0, n0) for _ in range(int(n))]
                qs = [uni10.Qnum(int(x)) for x in ((1, -1) if d == 2 else (2, 0, -2))]
                bA = [uni10.Bond(uni10.BD_IN, qa), uni10.Bond(uni10.BD_IN, qs), uni10.Bond(uni10.BD_OUT, qm)]
                bB = [uni10.Bond(uni10.BD_IN, qm), uni10.Bond(uni10.BD_IN, qs), uni10.Bond(uni10.BD_OUT, qa)]
            else:
                bA = [uni10.Bond(uni10.BD_IN, c), uni10.Bond(uni10.BD_IN, d), uni10.Bond(uni10.BD_OUT, c)]
                bB = bA
            A = uni10.UniTensor(bA)
            B = uni10.UniTensor(bB)
            rng = np.random.default_rng(r.seed)
            A.setRawElem(rng.uniform(-1, 1, c * d * c))
            B.setRawElem(rng.uniform(-1, 1, c * d * c))
            A.setLabel([0, 1, 2])
            B.setLabel([2, 3, 4])
            C = uni10.contract(A, B)
            return np.array([C.norm()])
        finally:
            uni10.set_gemm_mode(0)

    def run(self, requests: List[Request], realtime: bool = True) -> Dict:
        """Serve the requests; returns latency and throughput figures."""
        pending = sorted(requests, key=lambda r: r.t_arrival)
        waiting: Dict[tuple, List[Request]] = {c: [] for c in self.classes}
        contracts: List[Request] = []
        done: List[Request] = []
        t0 = time.perf_counter()
        i = 0
        clock = 0.0
        inflight = set()       # classes with a launched step not yet retired
        while i < len(pending) or contracts or any(waiting.values()) or any(self.busy.values()):
            now = (time.perf_counter() - t0) if realtime else clock
            if not realtime and i < len(pending) and not any(self.busy.values()) and not contracts \
                    and not any(waiting.values()):
                now = clock = max(clock, pending[i].t_arrival)
            while i < len(pending) and pending[i].t_arrival <= now:
                r = pending[i]
                (contracts if r.kind == "contract" else waiting[(r.chi, r.d, r.sym)]).append(r)
                i += 1
            # admission: free slots of classes between steps, first come first
            # served per class
            for c in self.classes:
                if c in inflight:
                    continue
                while waiting[c] and self.free[c]:
                    r = waiting[c].pop(0)
                    s = self.free[c].pop(0)
                    self.engines[c].load(s, r.delta, r.seed)
                    r.t_admit = now
                    self.busy[c][s] = [r, r.steps]
            # contractions run between steps
            for r in contracts:
                r.t_admit = now
                r.result = self._contract(r)
                r.t_done = (time.perf_counter() - t0) if realtime else now
                done.append(r)
            contracts = []
            # every class steps at its own pace: a class with busy slots and
            # no step in flight launches one batched step on its own stream
            # (the classes' steps overlap on the device), and a class whose
            # step has finished is retired without waiting for slower ones
            for c in self.classes:
                if self.busy[c] and c not in inflight:
                    self.engines[c].launch(1)
                    inflight.add(c)
            retired = False
            for c in self.classes:
                if c not in inflight or not self.engines[c].ready():
                    continue
                self.engines[c].wait()
                inflight.discard(c)
                retired = True
                t_done = (time.perf_counter() - t0) if realtime else now
                self.chain_steps += len(self.busy[c])
                for s in list(self.busy[c]):
                    ent = self.busy[c][s]
                    ent[1] -= 1
                    if ent[1] == 0:
                        r = ent[0]
                        r.result = np.concatenate([self.engines[c].lam(s, 0), self.engines[c].lam(s, 1)])
                        r.t_done = t_done
                        done.append(r)
                        del self.busy[c][s]
                        self.free[c].append(s)
            if not realtime:
                if retired:
                    clock = now + 1e-3
            elif i < len(pending) and not any(self.busy.values()) and not contracts \
                    and not any(waiting.values()):
                time.sleep(max(0.0, min(0.01, pending[i].t_arrival - (time.perf_counter() - t0))))
        wall = time.perf_counter() - t0
        lat = np.array([r.t_done - r.t_arrival for r in done])
        tebd_steps = sum(r.steps for r in done if r.kind == "tebd")
        return {
            "requests": len(done),
            "wall_s": wall,
            "p50_ms": float(np.percentile(lat, 50) * 1e3) if len(lat) else None,
            "p99_ms": float(np.percentile(lat, 99) * 1e3) if len(lat) else None,
            "tebd_steps_per_s": tebd_steps / wall,
            "chain_steps": self.chain_steps,
            "done": done,
        }
