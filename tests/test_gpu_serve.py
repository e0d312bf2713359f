"""Continuous-batching request loop (run with -m gpu): every result is the
same alone and in a mix of requests that arrive while others run, and the
TEBD requests follow the oracle's single-chain iTEBD."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def serve(tn):
    from trim_amd import serve as s
    from trim_amd import uni10
    uni10.load()
    return s


def _mix(serve):
    reqs = serve.poisson_requests(24, rate=4000.0, seed=5, chi_classes=[8, 16], steps_max=6)
    return reqs


def test_result_same_alone_and_in_mix(serve, golden):
    reqs = _mix(serve)
    budget = 4 * serve.slot_bytes(16) * 2          # about 4 slots per class: requests queue
    srv = serve.Server([8, 16], budget, golden)
    assert srv.slots(16) >= 1 and srv.slots(16) < 12
    stats = srv.run(reqs, realtime=False)
    _say("mix: %d requests, p50 %.2f ms p99 %.2f ms" % (stats["requests"], stats["p50_ms"], stats["p99_ms"]))
    assert stats["requests"] == len(reqs)
    mixed = {r.rid: r.result for r in stats["done"]}
    for target in (reqs[3], reqs[10], reqs[17]):
        alone = serve.Server([8, 16], budget, golden)
        solo = serve.poisson_requests(24, rate=4000.0, seed=5, chi_classes=[8, 16], steps_max=6)[target.rid]
        out = alone.run([solo], realtime=False)["done"][0].result
        _say("request %d (%s chi=%d steps=%d)" % (target.rid, target.kind, target.chi, target.steps))
        assert out.tobytes() == mixed[target.rid].tobytes()


def test_tebd_request_follows_oracle(serve, ora, golden):
    reqs = [serve.Request(rid=0, t_arrival=0.0, kind="tebd", chi=8, delta=0.8, seed=77, steps=5)]
    srv = serve.Server([8], 6 * serve.slot_bytes(8), golden)
    r = srv.run(reqs, realtime=False)["done"][0]
    O = ora.ITEBD(8, 2, ora.heisenberg_two_site(0.8), 0.01, 77, golden)
    for _ in range(5):
        O.step()
    assert np.max(np.abs(r.result[:8] - O.La)) <= 1e-10
    assert np.max(np.abs(r.result[8:] - O.Lb)) <= 1e-10


def test_mixed_classes_same_alone_and_in_mix(serve, ora, golden):
    """Requests of spin 1/2 and spin 1, dense and U(1) (60 %), TEBD steps and
    contractions, arriving while others run: each result is the same alone
    and in the mix, and a U(1) spin-1 TEBD request follows the oracle."""
    classes = [(c, d, s) for c in (8, 16) for d in (2, 3) for s in (0, 1)]
    gen = dict(n=32, rate=4000.0, seed=9, chi_classes=[8, 16], steps_max=5, d_choices=(2, 3), u1_fraction=0.6)
    reqs = serve.poisson_requests(**gen)
    assert {(r.d, r.sym) for r in reqs} == {(2, 0), (2, 1), (3, 0), (3, 1)}
    budget = 3 * serve.slot_bytes(16, 3) * len(classes)
    srv = serve.Server(classes, budget, golden)
    stats = srv.run(reqs, realtime=False)
    _say("mixed classes: %d requests, p50 %.2f ms p99 %.2f ms" % (stats["requests"], stats["p50_ms"], stats["p99_ms"]))
    assert stats["requests"] == len(reqs)
    mixed = {r.rid: r.result for r in stats["done"]}
    picks = [next(r for r in reqs if r.kind == k and r.sym == 1 and r.d == d) for k in ("tebd", "contract")
             for d in (2, 3)]
    for target in picks:
        solo = serve.poisson_requests(**gen)[target.rid]
        r1 = serve.Server(classes, budget, golden).run([solo], realtime=False)["done"][0]
        _say("request %d (%s chi=%d d=%d sym=%d)" % (target.rid, target.kind, target.chi, target.d, target.sym))
        assert r1.error is None, r1.error
        assert r1.result.tobytes() == mixed[target.rid].tobytes()
    t = next(r for r in picks if r.kind == "tebd" and r.d == 3)
    O = ora.ITEBDU1(t.chi, 3, t.delta, 0.01, t.seed, golden)
    for _ in range(t.steps):
        O.step()
    res = mixed[t.rid]
    assert np.max(np.abs(res[:t.chi] - O.lam[0])) <= 1e-10
    assert np.max(np.abs(res[t.chi:] - O.lam[1])) <= 1e-10


def test_configured_chi_range_same_alone_and_in_mix(serve, ora, golden):
    """Serving inside the configured range chi in [32, 1024]: classes chi 32
    (spin 1/2 dense), 64 (spin 1 U(1)), 256 (spin 1 U(1)) and 1024 (spin 1/2
    U(1)), TEBD and contraction requests arriving while others run.  Each
    picked result is the same alone and in the mix, and the chi = 256 spin-1
    U(1) TEBD request follows the oracle's per-charge-block update to 1e-10."""
    classes = [(32, 2, 0), (64, 3, 1), (256, 3, 1), (1024, 2, 1)]
    reqs = []
    rng = np.random.default_rng(12)
    t = 0.0
    for i in range(24):
        c = classes[i % len(classes)]
        t += float(rng.exponential(1.0 / 2000.0))
        kind = "tebd" if i % 3 != 2 else "contract"
        reqs.append(serve.Request(rid=i, t_arrival=t, kind=kind, chi=c[0], delta=float(rng.uniform(0, 2)),
                                  seed=500 + i, steps=int(rng.integers(1, 4)), d=c[1], sym=c[2]))
    budget = sum(2 * serve.slot_bytes(*c) for c in classes)
    srv = serve.Server(classes, budget * len(classes), golden, max_slots=2)
    stats = srv.run(reqs, realtime=False)
    _say("configured range: %d requests, p50 %.2f ms p99 %.2f ms" % (stats["requests"], stats["p50_ms"],
                                                                      stats["p99_ms"]))
    assert stats["requests"] == len(reqs)
    mixed = {r.rid: r.result for r in stats["done"]}
    picks = [r for r in reqs if r.chi in (256, 1024) and r.kind == "tebd"][:2] + \
            [next(r for r in reqs if r.kind == "contract" and r.sym == 1)]
    for target in picks:
        solo = dataclass_copy(target)
        out = serve.Server(classes, budget * len(classes), golden, max_slots=2).run([solo], realtime=False)
        _say("request %d (%s chi=%d d=%d sym=%d)" % (target.rid, target.kind, target.chi, target.d, target.sym))
        assert out["done"][0].result.tobytes() == mixed[target.rid].tobytes()
    t = next(r for r in reqs if r.chi == 256 and r.kind == "tebd")
    O = ora.ITEBDU1(256, 3, t.delta, 0.01, t.seed, golden)
    for _ in range(t.steps):
        O.step()
    res = mixed[t.rid]
    assert np.max(np.abs(res[:256] - O.lam[0])) <= 1e-10
    assert np.max(np.abs(res[256:] - O.lam[1])) <= 1e-10


def dataclass_copy(r):
    import dataclasses
    return dataclasses.replace(r, t_admit=-1.0, t_done=-1.0, result=None)


def _burst(serve, n, chi, steps=3, kind="tebd", seed0=700):
    return [serve.Request(rid=i, t_arrival=0.0, kind=kind, chi=chi, delta=0.3 + 0.1 * i, seed=seed0 + i,
                          steps=steps) for i in range(n)]


def test_over_budget_requests_queue(serve, tn, golden):
    """Eight requests at once against a budget of about two slots: they
    wait for slots instead of failing, the class keeps one engine, the
    library never holds more than the budget, and every result equals the
    one from a server with room to grow (four engines of two slots)."""
    import ctypes
    reqs = _burst(serve, 8, 16)
    tight = serve.Server([16], 3.5 * serve.slot_bytes(16), golden, slots_per_engine=2)
    try:
        assert tight.engine_count(16) == 1 and tight.slots(16) == 2
        tn.call("tn_pool_trim")
        stats = tight.run(reqs, realtime=False)
        v = [ctypes.c_size_t() for _ in range(4)]
        tn.call("tn_mem_stats", *[ctypes.byref(x) for x in v])
        assert v[0].value <= tight.budget
    finally:
        tight.close()
    _say("tight: %d requests, %d failed, engines %s" % (stats["requests"], stats["failed"], stats["engines"]))
    assert stats["requests"] == 8 and stats["failed"] == 0
    assert stats["engines"] == {"chi16_d2_dense": 1}
    roomy = serve.Server([16], 64 * serve.slot_bytes(16), golden, slots_per_engine=2)
    try:
        out = roomy.run(_burst(serve, 8, 16), realtime=False)
    finally:
        roomy.close()
    _say("roomy: engines %s" % out["engines"])
    assert out["engines"]["chi16_d2_dense"] >= 2
    a = {r.rid: r.result for r in stats["done"]}
    for r in out["done"]:
        assert r.result.tobytes() == a[r.rid].tobytes()


def test_budget_applies_only_to_server_allocations(serve, tn, golden):
    """The server's budget limits its own allocations only: outside its
    calls the library keeps the budget it had."""
    import ctypes
    b = ctypes.c_size_t()
    tn.call("tn_set_budget", 0)
    srv = serve.Server([8], 6 * serve.slot_bytes(8), golden)
    srv.run(_burst(serve, 2, 8, steps=2) + _burst(serve, 1, 8, kind="contract", seed0=950), realtime=False)
    tn.call("tn_mem_stats", None, None, None, ctypes.byref(b))
    assert b.value == 0 and srv.budget > 0
    srv.close()


def test_failed_engine_is_isolated(serve, tn, golden):
    """A class whose SVDs cannot converge (its plans capped at one sweep)
    fails its own requests with an error; the other class's requests and
    the contractions in the same stream complete, each equal to its result
    from a server without the failing class."""
    # max_slots=2: the failing class cannot grow an engine with default plans
    srv = serve.Server([8], 8 * serve.slot_bytes(16), golden, slots_per_engine=2, max_slots=2)
    tn.call("tn_svd_set_max_sweeps", 1)
    try:
        srv.add_class(16)
    finally:
        tn.call("tn_svd_set_max_sweeps", 0)
    reqs = _burst(serve, 3, 8) + [dataclass_copy(r) for r in _burst(serve, 3, 16, seed0=800)] \
        + _burst(serve, 2, 8, kind="contract", seed0=900)
    for i, r in enumerate(reqs):
        r.rid = i
    try:
        stats = srv.run(reqs, realtime=False)
    finally:
        srv.close()
    _say("isolation: %d requests, %d failed" % (stats["requests"], stats["failed"]))
    assert stats["requests"] == len(reqs)
    by = {r.rid: r for r in stats["done"]}
    for r in reqs:
        if r.chi == 16:
            assert by[r.rid].error is not None and "did not converge" in by[r.rid].error
        else:
            assert by[r.rid].error is None
    good = serve.Server([8], 8 * serve.slot_bytes(16), golden, slots_per_engine=2, max_slots=2)
    try:
        ref = good.run([dataclass_copy(r) for r in reqs if r.chi == 8], realtime=False)
    finally:
        good.close()
    for r in ref["done"]:
        assert r.result.tobytes() == by[r.rid].result.tobytes()


def test_idle_slots_are_switched_off(serve, golden):
    """Between requests every slot is inactive: a step of an engine holding
    one request computes one chain, and an idle server launches nothing."""
    srv = serve.Server([8], 12 * serve.slot_bytes(8), golden, slots_per_engine=4)
    try:
        e = srv.engines[(8, 2, 0)][0]
        assert not e.busy and not e.inflight
        stats = srv.run(_burst(serve, 1, 8, steps=4), realtime=False)
        assert stats["chain_steps"] == 4
        assert not e.busy and sorted(e.free[(8, 2, 0)]) == [0, 1, 2, 3]
    finally:
        srv.close()
