"""Continuous-batching request loop over the batched iTEBD engine.

Requests arrive on a Poisson clock and are of two kinds:

* ``tebd``: run ``steps`` iTEBD steps of one XXZ chain (bond dimension
  ``chi``, anisotropy ``delta``, seeded start) and return its Schmidt
  spectra;
* ``contract``: contract two seeded (chi, d, chi) tensors over one bond
  through the drop-in ``contract`` and return the result's norm.

A request also names its physical dimension d (2: spin 1/2, 3: spin 1) and
symmetry (dense, or U(1) with conserved Sz: block-sparse tensors whose
sectors follow the U(1) chains' sector rule).

TEBD requests run in batched engines, one kind per class (bond dimension, d,
symmetry), each engine a fixed number of slots whose tensors stay resident in
HBM.  A class starts with one engine and gains another when a request finds
every slot busy and the device memory budget (``tn_set_budget``, set by the
server) still admits one; when it does not, the request waits for a slot.
Every pass of the loop:

* admits arrived requests into free slots, first come first served per
  class; an admitted slot is restarted (identity warm-start bases) and
  switched on, a finished one switched off, so a step only computes busy
  slots (inactive chains are skipped on the device);
* launches one batched, graph-replayed step on every engine with busy slots
  and no step in flight, each on its own stream, so engines overlap on the
  device and fast ones do not wait for slow ones;
* retires every engine whose step has finished (a stream query that never
  blocks); an engine whose step reports a non-converged SVD fails the
  requests it holds -- they come back with an error, the loop goes on;
* issues contraction requests on the server's contraction stream: operands
  generated on the device from their seeds, the contraction and the norm
  enqueued without waiting; they are retired when that stream is idle.

A chain's computation never depends on which other slots are busy, so each
result is the same alone and in the mix.
"""
from __future__ import annotations

import contextlib
import ctypes
import dataclasses
import math
import time
from typing import Dict, List, Optional

import numpy as np

from . import tn, uni10


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
    error: Optional[str] = None


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
    """Upper bound of the HBM bytes one chain of the batched engine keeps
    resident (a dense chain's; a U(1) chain holds less).  For sizing
    budgets in tests and examples; admission itself asks the library."""
    D = chi * d
    doubles = 2 * chi * chi * d + 2 * chi + d ** 4 + 3 * chi * chi * d * d + 3 * D * D + D + 4
    return 8 * doubles


def _cls(c):
    """A class key (chi, d, sym); a bare chi means (chi, 2, 0)."""
    return (int(c), 2, 0) if isinstance(c, (int, np.integer)) else tuple(int(x) for x in c)


def _mem_stats():
    v = [ctypes.c_size_t() for _ in range(4)]
    tn.call("tn_mem_stats", *[ctypes.byref(x) for x in v])
    return [x.value for x in v]


class _Engine:
    """One batched engine of a class: its slots and the requests in them."""

    def __init__(self, cls, k, netdir, gemm_mode, tau):
        chi, d, sym = cls
        self.cls = cls
        self.eng = uni10.BatchTebd([chi] * k, [1.0] * k, [1 + s for s in range(k)], tau, netdir, gemm_mode,
                                   d=[d] * k, sym=[sym] * k)
        self.eng.set_graph(True)
        # the first (from-scratch) step and the graph capture, before any
        # request: its chains are placeholders, restarted on admission, so a
        # placeholder's unconverged SVD does not matter
        try:
            self.eng.step(2)
        except uni10.Uni10Error as e:
            if "did not converge" not in str(e):
                raise
        for s in range(k):
            self.eng.set_active(s, False)
        self.free = list(range(k))
        self.busy: Dict[int, list] = {}
        self.inflight = False


class Server:
    def __init__(self, classes, hbm_budget: float, netdir: Optional[str] = None, gemm_mode: int = 0,
                 max_slots: int = 256, tau: float = 0.01, slots_per_engine: Optional[int] = None):
        """classes: (chi, d, sym) tuples (or bare chi for (chi, 2, dense)).
        hbm_budget: the device memory the server may take beyond what the
        library holds when it starts; it is the library's budget
        (tn_set_budget) while the server allocates.  Every class starts with one engine
        sized to an equal share of it (slot_bytes per slot) and grows by
        engines of slots_per_engine slots (None: 64) while the budget
        admits them, up to max_slots slots per class."""
        self.netdir = netdir
        self.gemm_mode = gemm_mode
        self.tau = tau
        self.max_slots = max_slots
        self.k = max(1, min(max_slots, slots_per_engine or 64))
        self.classes = sorted(_cls(c) for c in classes)
        self.budget = int(_mem_stats()[0] + hbm_budget)
        self.engines: Dict[tuple, List[_Engine]] = {c: [] for c in self.classes}
        self.blocked: Dict[tuple, bool] = {c: False for c in self.classes}
        self.chain_steps = 0
        self.cstream = tn.stream_create()
        try:
            share = hbm_budget / len(self.classes)
            for c in self.classes:
                k0 = int(max(1, min(self.k, share // slot_bytes(*c))))
                if not self._grow(c, k0, limit=int(_mem_stats()[0] + share)):
                    raise uni10.Uni10Error("the device memory budget does not hold one engine of class %s" % (c,))
        except Exception:
            self.close()
            raise

    def close(self):
        """Release the engines (their memory returns to the pool)."""
        self.engines = {c: [] for c in self.classes}

    @contextlib.contextmanager
    def _budgeted(self, limit=None):
        """The server's budget (or a lower limit) as the library's while the
        server allocates (engines, contraction operands); the previous one
        afterwards, so other users of the library in the process are not
        limited by it."""
        prev = _mem_stats()[3]
        tn.call("tn_set_budget", self.budget if limit is None else min(limit, self.budget))
        try:
            yield
        finally:
            tn.call("tn_set_budget", int(prev))

    def add_class(self, cls):
        """Serve another class from now on (one engine created now)."""
        c = _cls(cls)
        if c in self.engines:
            return
        self.classes = sorted(self.classes + [c])
        self.engines[c] = []
        self.blocked[c] = False
        if not self._grow(c):
            raise uni10.Uni10Error("the device memory budget does not hold one engine of class %s" % (c,))

    def _grow(self, c, k=None, limit=None) -> bool:
        """One more engine for class c, of up to k slots (None: the
        engine size), as large as the budget left (or the lower limit on the
        library's memory in use) admits; False when not even one slot fits."""
        # pool blocks are reusable as soon as they are freed: contractions
        # still queued may use blocks that the new engine would take
        tn.call("tn_stream_sync", self.cstream)
        have = sum(len(e.eng.chi) for e in self.engines[c])
        k = min(k or self.k, self.max_slots - have)
        left = (self.budget if limit is None else min(limit, self.budget)) - _mem_stats()[0]
        k = min(k, int(left // slot_bytes(*c)))
        while k >= 1:
            try:
                with self._budgeted(limit):
                    self.engines[c].append(_Engine(c, k, self.netdir, self.gemm_mode, self.tau))
                return True
            except uni10.Uni10Error as e:
                if "budget exceeded" not in str(e):
                    raise
                k //= 2           # plans and workspaces: the estimate was short
        return False

    def warmup(self):
        """Engines are warmed when created; kept for callers of the old API."""

    def slots(self, cls) -> int:
        return sum(len(e.eng.chi) for e in self.engines[_cls(cls)])

    def engine_count(self, cls) -> int:
        return len(self.engines[_cls(cls)])

    # -- contraction requests -------------------------------------------------
    def _contract_bonds(self, c, d, sym):
        if sym:
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
        return bA, bB

    def _issue_contract(self, r: Request):
        """Enqueue one contraction on the contraction stream: operands seeded
        on the device, contract, norm into a device scalar (read at
        retirement).  Nothing here waits for the device."""
        uni10.set_gemm_mode(self.gemm_mode)
        uni10.set_stream(self.cstream)
        try:
            with self._budgeted():
                bA, bB = self._contract_bonds(r.chi, r.d, r.sym)
                A = uni10.UniTensor(bA).randomize(r.seed)
                B = uni10.UniTensor(bB).randomize(r.seed + 1)
                A.setLabel([0, 1, 2])
                B.setLabel([2, 3, 4])
                C = uni10.contract(A, B)
                out = tn.DeviceArray((1,))
                tn.call("tn_norm2_f64", ctypes.c_void_p(C.device_ptr()), C.elemNum(), tn._p(out), self.cstream)
            return (A, B, C, out)
        finally:
            uni10.set_stream(None)
            uni10.set_gemm_mode(0)

    def run(self, requests: List[Request], realtime: bool = True) -> Dict:
        """Serve the requests; returns latency and throughput figures."""
        pending = sorted(requests, key=lambda r: r.t_arrival)
        waiting: Dict[tuple, List[Request]] = {c: [] for c in self.classes}
        contracts: List[tuple] = []          # (request, resident tensors) in flight
        cwait: List[Request] = []            # contractions the budget holds back
        done: List[Request] = []
        t0 = time.perf_counter()
        i = 0
        clock = 0.0

        def now_s():
            return (time.perf_counter() - t0) if realtime else clock

        def busy_any():
            return any(e.busy for es in self.engines.values() for e in es)

        while i < len(pending) or contracts or cwait or any(waiting.values()) or busy_any():
            now = now_s()
            if not realtime and i < len(pending) and not busy_any() and not contracts and not cwait \
                    and not any(waiting.values()):
                now = clock = max(clock, pending[i].t_arrival)
            while i < len(pending) and pending[i].t_arrival <= now:
                r = pending[i]
                i += 1
                if r.kind == "contract":
                    cwait.append(r)
                else:
                    waiting[(r.chi, r.d, r.sym)].append(r)
            # contractions are issued while the budget admits their tensors;
            # one that does not fit even with nothing else in flight fails
            while cwait:
                r = cwait[0]
                try:
                    contracts.append((r, self._issue_contract(r)))
                    r.t_admit = now
                    cwait.pop(0)
                except (uni10.Uni10Error, tn.TNError) as err:
                    if "budget exceeded" not in str(err):
                        raise
                    if contracts or busy_any():
                        break
                    r.error = str(err)
                    r.t_done = now
                    done.append(cwait.pop(0))
            # admission: free slots of engines between steps; a class whose
            # slots are all busy grows while the memory budget admits it
            for c in self.classes:
                if not waiting[c]:
                    continue
                for e in self.engines[c]:
                    if e.inflight:
                        continue
                    while waiting[c] and e.free:
                        r = waiting[c].pop(0)
                        s = e.free.pop(0)
                        e.eng.load(s, r.delta, r.seed)
                        e.eng.set_active(s, True)
                        r.t_admit = now
                        e.busy[s] = [r, r.steps]
                if waiting[c] and not self.blocked[c] and all(not e.free for e in self.engines[c]):
                    if not self._grow(c):
                        self.blocked[c] = True      # until one of its slots frees
            # every engine with busy slots and no step in flight launches one step
            for es in self.engines.values():
                for e in es:
                    if e.busy and not e.inflight:
                        e.eng.launch(1)
                        e.inflight = True
            retired = False
            for c, es in self.engines.items():
                for e in es:
                    if not e.inflight or not e.eng.ready():
                        continue
                    e.inflight = False
                    retired = True
                    t_done = now_s() if realtime else now
                    try:
                        e.eng.wait()
                        failed = None
                    except uni10.Uni10Error as err:
                        failed = str(err)
                    self.chain_steps += len(e.busy)
                    for s in list(e.busy):
                        ent = e.busy[s]
                        ent[1] -= 1
                        if ent[1] == 0 or failed:
                            r = ent[0]
                            if failed:
                                r.error = failed
                            else:
                                r.result = np.concatenate([e.eng.lam(s, 0), e.eng.lam(s, 1)])
                            r.t_done = t_done
                            done.append(r)
                            del e.busy[s]
                            e.eng.set_active(s, False)
                            e.free.append(s)
                            self.blocked[c] = False
            if contracts:
                ok = ctypes.c_int()
                tn.call("tn_stream_query", self.cstream, ctypes.byref(ok))
                if ok.value:
                    t_done = now_s() if realtime else now
                    for r, keep in contracts:
                        r.result = keep[3].to_host(self.cstream)
                        r.t_done = t_done
                        done.append(r)
                    contracts = []
                    retired = True
            if not realtime:
                if retired:
                    clock = now + 1e-3
            elif i < len(pending) and not busy_any() and not contracts and not cwait \
                    and not any(waiting.values()):
                time.sleep(max(0.0, min(0.01, pending[i].t_arrival - (time.perf_counter() - t0))))
        wall = time.perf_counter() - t0
        ok_done = [r for r in done if r.error is None]
        lat = np.array([r.t_done - r.t_arrival for r in ok_done])
        tebd_steps = sum(r.steps for r in ok_done if r.kind == "tebd")
        return {
            "requests": len(done),
            "failed": len(done) - len(ok_done),
            "wall_s": wall,
            "requests_per_s": len(done) / wall if wall > 0 else None,
            "p50_ms": float(np.percentile(lat, 50) * 1e3) if len(lat) else None,
            "p99_ms": float(np.percentile(lat, 99) * 1e3) if len(lat) else None,
            "tebd_steps_per_s": tebd_steps / wall,
            "chain_steps": self.chain_steps,
            "engines": {"chi%d_d%d_%s" % (c[0], c[1], "u1" if c[2] else "dense"): len(es)
                        for c, es in self.engines.items()},
            "done": done,
        }
