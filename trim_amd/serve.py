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

TEBD requests run in batched engines whose slots hold chains resident in
HBM.  A class (bond dimension, d, symmetry) is probed once for the time of
one chain's step and sorted into a light, medium or heavy lane; the classes
of a lane share one engine, so a step of that engine is one grouped launch
per stage over the busy slots of all of them.  The device runs the work of
one stream at a time per hardware queue, so the server keeps four streams:
the three lanes and one for contractions.  A lane runs one engine step at a
time, its engines taking turns, so a bond-dimension-1024 step never sits in
front of a bond-dimension-32 one.  Memory follows a budget (the library's
``tn_set_budget`` while the server allocates): the classes get equal shares
for their first slots, and a light class whose slots are all busy gains an
engine of its own while the budget admits one; otherwise its requests wait
for a slot.  Every pass of the loop:

* admits arrived requests into free slots, first come first served per
  class; an admitted slot is restarted (identity warm-start bases) and
  switched on, a finished one switched off, so a step only computes busy
  slots (inactive chains are skipped on the device);
* on every idle lane, launches one batched, graph-replayed step of the next
  engine with busy slots;
* retires the engine of every lane whose step has finished (a stream query
  that never blocks); an engine whose step reports a non-converged SVD
  fails the requests it holds -- they come back with an error, the loop
  goes on;
* issues contraction requests (at most 32 on the device at once) on the
  contraction stream: operands generated
  on the device from their seeds, the contraction, its norm and the copy of
  the norm to pinned host memory enqueued without waiting; each is retired
  when the event recorded after it has completed.

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

# lanes by the time of one chain's warm step (milliseconds)
LIGHT_MS = 8.0
HEAVY_MS = 60.0
# classes whose chain step is at least this long gain no second engine
GROW_MAX_MS = 20.0
# per lane, the most one-chain step time (ms) the busy slots may add up to:
# a lane's step stays short enough not to hold the device from the others
# (a lane with no busy slot always admits one request)
LANE_MS = (float("inf"), 400.0, 2500.0)
# contraction results in flight at once (pinned host slots)
MAX_CONTRACTS = 4096
# engines per class: a lane runs its engines in turn, so more engines add
# slots but no speed
MAX_ENGINES_PER_CLASS = 4
# part of the memory budget that engines leave to contractions
CONTRACT_RESERVE = 0.1
# contractions on the device at once (later ones wait on the host)
MAX_INFLIGHT_CONTRACTS = 32


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


def _probe_ms(cls, netdir, gemm_mode, tau) -> float:
    """Time of one warm step of one chain of class cls (a throwaway
    one-chain engine: a cold step, then the timed one); inf when its SVD
    does not converge."""
    chi, d, sym = cls
    eng = uni10.BatchTebd([chi], [1.0], [1], tau, netdir, gemm_mode, d=[d], sym=[sym])
    try:
        eng.step(1)
        t = time.perf_counter()
        eng.step(1)
        return (time.perf_counter() - t) * 1e3
    except uni10.Uni10Error as e:
        if "did not converge" not in str(e):
            raise
        return float("inf")
    finally:
        del eng


class _Engine:
    """One batched engine on a lane: slots of one or more classes (one
    grouped launch per stage covers the busy slots of all of them) and the
    requests in them."""

    def __init__(self, spec, netdir, gemm_mode, tau, lane, warm):
        """spec: [(class, slots), ...]."""
        chi, d, sym, self.cls_of = [], [], [], []
        self.free: Dict[tuple, List[int]] = {}
        for c, k in spec:
            self.free[c] = list(range(len(chi), len(chi) + k))
            chi += [c[0]] * k
            d += [c[1]] * k
            sym += [c[2]] * k
            self.cls_of += [c] * k
        n = len(chi)
        self.eng = uni10.BatchTebd(chi, [1.0] * n, [1 + s for s in range(n)], tau, netdir, gemm_mode, d=d, sym=sym)
        self.eng.set_graph(True)
        for s in range(n):
            self.eng.set_active(s, False)
        if warm:
            # the from-scratch step and the graph capture, with every slot
            # off (every stage skips them), before any request
            self.eng.step(2)
        self.eng.set_stream(lane)
        self.busy: Dict[int, list] = {}
        self.inflight = False

    def slots(self, c) -> int:
        return self.cls_of.count(c)


class Server:
    def __init__(self, classes, hbm_budget: float, netdir: Optional[str] = None, gemm_mode: int = 0,
                 max_slots: int = 256, tau: float = 0.01, slots_per_engine: Optional[int] = None,
                 step_ms: Optional[Dict[tuple, float]] = None):
        """classes: (chi, d, sym) tuples (or bare chi for (chi, 2, dense)).
        hbm_budget: the device memory the server may take beyond what the
        library holds when it starts; it is the library's budget
        (tn_set_budget) while the server allocates.  Engines take at most
        90 % of it (the rest is left to contractions): every class gets an
        equal share of that for its first slots (slot_bytes per slot, at most
        slots_per_engine, None: 64); the classes of a lane share one engine.
        A class whose slots are all busy gains an engine of its own while
        the budget admits one, up to max_slots slots.  step_ms: the classes'
        one-chain step times from an earlier server (skips the probes)."""
        self.netdir = netdir
        self.gemm_mode = gemm_mode
        self.tau = tau
        self.max_slots = max_slots
        self.k = max(1, min(max_slots, slots_per_engine or 64))
        self.classes: List[tuple] = []
        self.budget = int(_mem_stats()[0] + hbm_budget)
        self.lane_engines: List[List[_Engine]] = [[], [], []]
        self.blocked: Dict[tuple, bool] = {}
        self.lane_of: Dict[tuple, int] = {}
        self.step_ms: Dict[tuple, float] = dict(step_ms or {})
        self.chain_steps = 0
        # four streams: light, medium and heavy engine lanes, contractions
        self.lanes = [tn.stream_create() for _ in range(3)]
        self.cstream = tn.stream_create()
        self._pinned = ctypes.c_void_p()
        tn.call("tn_host_alloc", ctypes.byref(self._pinned), 8 * MAX_CONTRACTS)
        self._pinned_view = (ctypes.c_double * MAX_CONTRACTS).from_address(self._pinned.value)
        self._free_pins = list(range(MAX_CONTRACTS))
        cl = sorted(set(_cls(c) for c in classes))
        # engines take at most (1 - CONTRACT_RESERVE) of the budget, so that
        # contractions always find room
        self._share = hbm_budget * (1.0 - CONTRACT_RESERVE) / max(1, len(cl))
        self._reserve = int(hbm_budget * CONTRACT_RESERVE)
        try:
            for c in cl:
                self._register(c)
            # light and medium classes share their lane's engine; heavy
            # classes get one each (a heavy step of one class does not wait
            # for a longer one of another)
            for ln in range(3):
                mine = [c for c in cl if self.lane_of[c] == ln]
                groups = [mine] if ln < 2 else [[c] for c in mine]
                for g in groups:
                    spec = [(c, self._first_slots(c)) for c in g]
                    if spec and not self._build(ln, spec, limit=int(_mem_stats()[0] + self._share * len(spec))):
                        raise uni10.Uni10Error("the device memory budget does not hold an engine of lane %d" % ln)
        except Exception:
            self.close()
            raise

    def _register(self, c):
        if c not in self.step_ms:
            self.step_ms[c] = _probe_ms(c, self.netdir, self.gemm_mode, self.tau)
        ms = self.step_ms[c]
        self.lane_of[c] = 0 if ms < LIGHT_MS else (1 if ms < HEAVY_MS else 2)
        self.classes = sorted(self.classes + [c])
        self.blocked[c] = False

    def _first_slots(self, c) -> int:
        """Slots of a class's first engine: its memory share, and no more
        than its lane's step limit lets be busy at once."""
        cap = LANE_MS[self.lane_of[c]] // max(self.step_ms[c], 1e-3)
        return int(max(1, min(self.k, self._share // slot_bytes(*c), cap)))

    @property
    def engines(self) -> Dict[tuple, List[_Engine]]:
        """The engines holding slots of each class."""
        return {c: [e for e in self.lane_engines[self.lane_of[c]] if e.slots(c)] for c in self.classes}

    def close(self):
        """Release the engines (their memory returns to the pool), the
        streams and the pinned result slots."""
        for es in getattr(self, "lane_engines", []):
            for e in es:
                e.eng.set_stream(None)
        self.lane_engines = [[], [], []]
        for h in getattr(self, "lanes", []) + [getattr(self, "cstream", None)]:
            if h:
                tn.call("tn_stream_sync", h)
                tn.call("tn_stream_destroy", h)
        self.lanes, self.cstream = [None, None, None], None
        if getattr(self, "_pinned", None):
            tn.call("tn_host_free", self._pinned)
            self._pinned = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

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
        """Serve another class from now on: its step time is probed, its
        lane chosen, and an engine for it created now."""
        c = _cls(cls)
        if c in self.lane_of:
            return
        self._register(c)
        if not self._build(self.lane_of[c], [(c, self._first_slots(c))],
                           limit=int(_mem_stats()[0] + self._share)):
            raise uni10.Uni10Error("the device memory budget does not hold one engine of class %s" % (c,))

    def _build(self, ln, spec, limit=None, warm=True) -> bool:
        """An engine with the given slots on lane ln, halving every class's
        slots while the budget (or the lower limit on the library's memory
        in use) refuses it; False when not even one slot each fits."""
        # pool blocks are reusable as soon as they are freed: contractions
        # still queued may use blocks that the new engine would take
        tn.call("tn_stream_sync", self.cstream)
        left = (self.budget if limit is None else min(limit, self.budget)) - _mem_stats()[0]
        need = sum(k * slot_bytes(*c) for c, k in spec)
        if need > left:
            f = max(0.0, left / need)
            spec = [(c, max(1, int(k * f))) for c, k in spec]
        while True:
            try:
                with self._budgeted(limit):
                    e = _Engine(spec, self.netdir, self.gemm_mode, self.tau, self.lanes[ln], warm)
                self.lane_engines[ln].append(e)
                return True
            except uni10.Uni10Error as err:
                if "budget exceeded" not in str(err) or all(k == 1 for _, k in spec):
                    if "budget exceeded" in str(err):
                        return False
                    raise
                # plans and workspaces: the estimate was short
                spec = [(c, max(1, k // 2)) for c, k in spec]

    def _grow(self, c) -> bool:
        """An engine of its own for class c (no warm-up: its first launch
        runs the from-scratch step on the lane), if the budget admits it."""
        mine = [e for e in self.lane_engines[self.lane_of[c]] if e.slots(c)]
        if len(mine) >= MAX_ENGINES_PER_CLASS:
            return False
        have = sum(e.slots(c) for e in mine)
        k = min(self.k, self.max_slots - have)
        return k >= 1 and self._build(self.lane_of[c], [(c, k)], limit=self.budget - self._reserve, warm=False)

    def warmup(self):
        """Engines are warmed when created; kept for callers of the old API."""

    def slots(self, cls) -> int:
        c = _cls(cls)
        return sum(e.slots(c) for e in self.lane_engines[self.lane_of[c]])

    def engine_count(self, cls) -> int:
        return len(self.engines[_cls(cls)])

    def lane(self, cls) -> int:
        """0 light, 1 medium, 2 heavy."""
        return self.lane_of[_cls(cls)]

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

    def _issue_contract(self, r: Request, pin: int):
        """Enqueue one contraction on the contraction stream: operands seeded
        on the device, contract, norm into a device scalar, its copy into
        pinned slot `pin`, an event after it.  Nothing here waits for the
        device."""
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
            tn.call("tn_memcpy_d2h_async", ctypes.c_void_p(self._pinned.value + 8 * pin), tn._p(out), 8,
                    self.cstream)
            ev = tn.Event()
            ev.record(self.cstream)
            return (A, B, C, out, ev, pin)
        finally:
            uni10.set_stream(None)
            uni10.set_gemm_mode(0)

    def run(self, requests: List[Request], realtime: bool = True, deadline_s: Optional[float] = None) -> Dict:
        """Serve the requests; returns latency and throughput figures.

        deadline_s (realtime only): a serving window.  At the deadline the
        loop stops taking arrivals, admitting and launching, lets the work
        on the device finish and returns; requests that arrived and did not
        complete are counted as unfinished (an overloaded window), requests
        due after it as not offered."""
        pending = sorted(requests, key=lambda r: r.t_arrival)
        waiting: Dict[tuple, List[Request]] = {c: [] for c in self.classes}
        contracts: List[tuple] = []          # (request, resident tensors, event, pin) in flight
        cwait: List[Request] = []            # contractions the budget or the pinned slots hold back
        done: List[Request] = []
        lane_busy: List[Optional[_Engine]] = [None, None, None]
        turn = [0, 0, 0]                     # round-robin position per lane
        lane_ms = [0.0, 0.0, 0.0]            # one-chain step times of the busy slots, per lane
        t0 = time.perf_counter()
        i = 0
        clock = 0.0
        steps0 = self.chain_steps
        host = {"grow": 0.0, "contract": 0.0, "launch": 0.0}     # host seconds per activity
        stopping = False

        def now_s():
            return (time.perf_counter() - t0) if realtime else clock

        def busy_any():
            return any(e.busy for es in self.lane_engines for e in es)

        def idle():
            return not busy_any() and not contracts and not cwait and not any(waiting.values())

        while i < len(pending) or contracts or cwait or any(waiting.values()) or busy_any():
            now = now_s()
            if deadline_s is not None and realtime and now >= deadline_s:
                stopping = True
            if stopping and not contracts and all(x is None for x in lane_busy):
                break
            if not realtime and i < len(pending) and idle():
                now = clock = max(clock, pending[i].t_arrival)
            while not stopping and i < len(pending) and pending[i].t_arrival <= now:
                r = pending[i]
                i += 1
                (cwait if r.kind == "contract" else waiting[(r.chi, r.d, r.sym)]).append(r)
            # contractions are issued while the budget admits their tensors
            # and a pinned result slot is free; one that does not fit even
            # with nothing else in flight fails
            while cwait and not stopping and self._free_pins and len(contracts) < MAX_INFLIGHT_CONTRACTS:
                r = cwait[0]
                tc = time.perf_counter()
                try:
                    contracts.append((r, self._issue_contract(r, self._free_pins[-1])))
                    self._free_pins.pop()
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
                finally:
                    host["contract"] += time.perf_counter() - tc
            # admission: free slots of engines between steps; a class whose
            # slots are all busy grows while the memory budget admits it and
            # its chain step is short (a long one would hold the loop)
            for c in self.classes:
                if not waiting[c] or stopping:
                    continue
                ln = self.lane_of[c]
                mine = [e for e in self.lane_engines[ln] if c in e.free]
                for e in mine:
                    if e.inflight:
                        continue
                    while waiting[c] and e.free[c] and \
                            (lane_ms[ln] == 0.0 or lane_ms[ln] + self.step_ms[c] <= LANE_MS[ln]):
                        lane_ms[ln] += self.step_ms[c]
                        r = waiting[c].pop(0)
                        s = e.free[c].pop(0)
                        e.eng.load(s, r.delta, r.seed)
                        e.eng.set_active(s, True)
                        r.t_admit = now
                        e.busy[s] = [r, r.steps]
                if waiting[c] and not self.blocked[c] and self.step_ms[c] < GROW_MAX_MS and \
                        all(not e.free[c] for e in mine):
                    tg = time.perf_counter()
                    grown = self._grow(c)
                    host["grow"] += time.perf_counter() - tg
                    if not grown:
                        self.blocked[c] = True      # until one of its slots frees
            # every idle lane launches one step of its next engine with busy
            # slots (round robin over the lane's engines)
            if not stopping:
                for ln in range(3):
                    if lane_busy[ln] is not None:
                        continue
                    cand = self.lane_engines[ln]
                    for k in range(len(cand)):
                        e = cand[(turn[ln] + k) % len(cand)]
                        if e.busy:
                            tl = time.perf_counter()
                            e.eng.launch(1)
                            host["launch"] += time.perf_counter() - tl
                            e.inflight = True
                            lane_busy[ln] = e
                            turn[ln] = (turn[ln] + k + 1) % len(cand)
                            break
            retired = False
            for ln in range(3):
                e = lane_busy[ln]
                if e is None or not e.eng.ready():
                    continue
                lane_busy[ln] = None
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
                        e.free[e.cls_of[s]].append(s)
                        self.blocked[e.cls_of[s]] = False
                        lane_ms[ln] = max(0.0, lane_ms[ln] - self.step_ms[e.cls_of[s]])
            if contracts:
                t_done = now_s() if realtime else now
                still = []
                for ent in contracts:
                    r, keep = ent
                    ev, pin = keep[4], keep[5]
                    if ev.query():
                        r.result = np.array([self._pinned_view[pin]])
                        r.t_done = t_done
                        done.append(r)
                        ev.destroy()
                        self._free_pins.append(pin)
                        retired = True
                    else:
                        still.append(ent)
                contracts = still
            if not realtime:
                if retired:
                    clock = now + 1e-3
            elif i < len(pending) and idle():
                time.sleep(max(0.0, min(0.01, pending[i].t_arrival - (time.perf_counter() - t0))))
        wall = time.perf_counter() - t0
        unfinished = len(cwait) + sum(len(w) for w in waiting.values()) + \
            sum(len(e.busy) for es in self.lane_engines for e in es)
        for es in self.lane_engines:          # slots of unfinished requests go back
            for e in es:
                for s in list(e.busy):
                    e.eng.set_active(s, False)
                    e.free[e.cls_of[s]].append(s)
                e.busy = {}
        ok_done = [r for r in done if r.error is None]
        lat = np.array([r.t_done - r.t_arrival for r in ok_done])
        tebd_steps = sum(r.steps for r in ok_done if r.kind == "tebd")
        return {
            "requests": len(done),
            "offered": i,
            "unfinished": unfinished,
            "failed": len(done) - len(ok_done),
            "wall_s": wall,
            "requests_per_s": len(done) / wall if wall > 0 else None,
            "p50_ms": float(np.percentile(lat, 50) * 1e3) if len(lat) else None,
            "p99_ms": float(np.percentile(lat, 99) * 1e3) if len(lat) else None,
            "tebd_steps_per_s": tebd_steps / wall,
            "chain_steps": self.chain_steps - steps0,
            "chain_steps_per_s": (self.chain_steps - steps0) / wall,
            "host_s": host,
            "engines": {"chi%d_d%d_%s" % (c[0], c[1], "u1" if c[2] else "dense"): len(es)
                        for c, es in self.engines.items()},
            "lanes": {"chi%d_d%d_%s" % (c[0], c[1], "u1" if c[2] else "dense"): self.lane_of[c]
                      for c in self.classes},
            "done": done,
        }
