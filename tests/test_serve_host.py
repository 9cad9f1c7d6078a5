"""Host logic of the serving loop (nec2c_amd/serve.py) on the CPU: arrival-order
admission under an HBM footprint budget, LRU eviction of idle cached systems,
cache hits on repeated (grid, frequency, eps) keys, rejection of a request that
can never fit, and the latency report.  The device batch is replaced by a
stand-in that finishes each job a fixed number of progress calls after
submission, so no GPU is touched."""
import types

import numpy as np
import pytest

from nec2c_amd import serve


class FakeSystem:
    def __init__(self, key):
        self.key = key
        self.closed = False

    def close(self):
        self.closed = True


class FakeBatch:
    def __init__(self, delay=3):
        self.delay = delay
        self.jobs = {}
        self.next = 1
        self.in_flight_max = 0

    def submit(self, system, b, **kw):
        jid = self.next
        self.next += 1
        self.jobs[jid] = [system, self.delay]
        self.in_flight_max = max(self.in_flight_max, len(self.jobs))
        return jid

    def progress(self):
        ready = []
        for jid, j in list(self.jobs.items()):
            j[1] -= 1
            if j[1] <= 0:
                ready.append(jid)
        return ready, len(self.jobs) - len(ready)

    def result(self, jid):
        self.jobs.pop(jid)
        rep = types.SimpleNamespace(relative_residual=5e-9, iterations=10)
        return True, np.zeros(1), rep, 0.001

    def close(self):
        pass


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 1e-4
        return self.t


def make_server(budget, delay=3):
    srv = serve.Server(budget, clock=Clock())
    srv.batch.close()
    srv.batch = FakeBatch(delay)
    return srv


def request(rid, key, n, ports, built, arrival=0.0):
    def make():
        built.append(key)
        return FakeSystem(key)
    return serve.Request(rid, arrival, key, (n, n, n), 21, make, [np.zeros(4)] * ports)


def test_footprint_model_is_monotone_and_positive():
    a = serve.system_device_bytes((16, 16, 16), 21)
    b = serve.system_device_bytes((32, 32, 32), 21)
    assert 0 < a < b
    assert serve.job_device_bytes((16, 16, 16), 21, 120) > serve.job_device_bytes((16, 16, 16), 21, 30)


def test_cache_hits_and_arrival_order():
    built = []
    srv = make_server(budget=10 ** 12)
    reqs = [request(0, "A", 16, 2, built), request(1, "B", 16, 1, built),
            request(2, "A", 16, 3, built, arrival=0.0)]
    rep = srv.run(reqs)
    assert rep["requests"] == 3 and rep["solves"] == 6
    assert built == ["A", "B"]          # the repeated key reuses the cached system
    assert rep["cache_hits"] == 1 and rep["system_builds"] == 2
    assert [r.rid for r in sorted(srv.done, key=lambda r: r.admitted)] == [0, 1, 2]
    assert rep["p50_ms"] is not None and rep["p99_ms"] >= rep["p50_ms"]


def test_budget_limits_concurrency_and_evicts_idle_systems():
    built = []
    n = 16
    sys_b = serve.system_device_bytes((n,) * 3, 21)
    one = sys_b + serve.job_device_bytes((n,) * 3, 21, 120)
    srv = make_server(budget=one + sys_b // 2)   # room for one model and one port at a time
    reqs = [request(i, f"K{i}", n, 1, built) for i in range(4)]
    rep = srv.run(reqs)
    assert rep["requests"] == 4 and rep["rejected"] == 0
    assert srv.batch.in_flight_max == 1      # never two requests' jobs on the device together
    assert len(srv.cache) == 1               # idle systems were evicted to make room
    assert srv.used <= srv.budget


def test_request_that_can_never_fit_is_rejected():
    built = []
    srv = make_server(budget=1000)
    rep = srv.run([request(0, "huge", 32, 1, built)])
    assert rep["rejected"] == 1 and rep["requests"] == 0
    assert built == []


@pytest.mark.parametrize("delay", [1, 5])
def test_run_batch_keeps_inflight_cap(delay, monkeypatch):
    fb = FakeBatch(delay)
    jobs = [(FakeSystem(i), np.zeros(2)) for i in range(10)]
    res, secs = serve.run_batch(jobs, inflight=3, batch=fb)
    assert len(res) == 10 and all(r[0] for r in res)
    assert fb.in_flight_max <= 3


def test_returning_key_that_is_lru_is_not_evicted_for_itself():
    """A request whose key is cached (and least recently used) must not evict
    its own system to make room: the footprint counted that system as resident,
    so evicting it would rebuild it and push the accounted bytes over budget."""
    built = []
    n = 16
    sys_b = serve.system_device_bytes((n,) * 3, 21)
    job_b = serve.job_device_bytes((n,) * 3, 21, 120)
    srv = make_server(budget=2 * sys_b + 2 * job_b - 1)   # K1 + K2 cached fit, plus one job
    rep = srv.run([request(0, "K1", n, 1, built), request(1, "K2", n, 1, built, arrival=0.01)])
    assert rep["requests"] == 2 and built == ["K1", "K2"]
    assert list(srv.cache) == ["K1", "K2"]           # K1 is the LRU entry now
    used_max = []
    step = srv.step
    srv.step = lambda arrived: (step(arrived), used_max.append(srv.used))[0]
    srv.run([request(2, "K1", n, 2, built)])         # two ports: K2 (idle) must go, K1 must stay
    assert built == ["K1", "K2"]                     # one build each: K1 came from the cache
    assert srv.done[-1].cache_hit
    assert max(used_max) <= srv.budget
    assert list(srv.cache) == ["K1"]


class RecordingBatch(FakeBatch):
    """FakeBatch that records submissions and collected results in order."""

    def __init__(self, delay=3):
        super().__init__(delay)
        self.log = []

    def submit(self, system, b, **kw):
        jid = super().submit(system, b, **kw)
        self.log.append(("submit", system.key, jid))
        return jid

    def result(self, jid):
        self.log.append(("result", None, jid))
        return super().result(jid)


def test_seed_port_runs_first_on_a_new_model():
    """On a model without a published subspace, a request's first port runs
    alone and the others are submitted when it finishes; a later request for
    the cached (seeded) model submits all its ports at once."""
    built = []
    srv = make_server(budget=10 ** 12)
    srv.batch = RecordingBatch(delay=3)
    reqs = [request(0, "A", 16, 3, built), request(1, "A", 16, 2, built, arrival=0.005)]
    rep = srv.run(reqs)
    assert rep["requests"] == 2 and rep["solves"] == 5 and built == ["A"]
    log = srv.batch.log
    # the seed (job 1) alone, its result, then ports 2 and 3 of the first request
    assert log[:4] == [("submit", "A", 1), ("result", None, 1), ("submit", "A", 2), ("submit", "A", 3)]
    # the second request's two ports go in together
    second = [e for e in log if e[0] == "submit" and e[2] >= 4]
    first_result_after = min(i for i, e in enumerate(log) if e[0] == "result" and e[2] >= 4)
    assert len(second) == 2 and log.index(second[1]) < first_result_after
    assert srv.used == sum(e[1] for e in srv.cache.values())   # only the cached model remains


def test_held_ports_count_against_the_budget_at_admission():
    built = []
    srv = make_server(budget=10 ** 12)
    srv.batch = RecordingBatch(delay=50)
    srv.t0 = srv.clock()
    req = request(0, "A", 16, 3, built)
    srv._admit(req)
    assert len(srv.inflight) == 1
    assert srv.used == srv._system_bytes(req) + 3 * srv._job_bytes(req)
