"""The serving loop (nec2c_amd/serve.py) on the device: Poisson arrivals of
frequency points for 16^3-32^3 random-permittivity models with key reuse,
1-3 ports each, mixed precision to 1e-8, admitted while others are in flight.

Every port's result must equal the same request solved alone (the same
system and options through a batch of one job) within 1e-12 relative, with
the same iteration count; every result must meet 1e-8 on the reference's own
operator (oracle/_ref); cached systems must be reused; and requests must
overlap on the device (a request admitted before an earlier one finished).
A second run under a budget that holds one model at a time must give the
same results while never holding more bytes than the budget.  These two runs
keep each job on its own (share=False), so the mix may not change a single
bit.  A third run shares each model's recycled subspace among its ports and
later requests (the server's default): there the results agree with the solo
solves to the solver tolerance (feed impedance within 1e-6, the reference
operator's residual within 1.05e-8) and need fewer Arnoldi steps."""
import math

import numpy as np
import pytest

from nec2c_amd import serve, vie, workloads as W

pytestmark = pytest.mark.gpu

PITCH = 0.5 / 64
OPTS = dict(restart=60, maximum_iterations=1500, relative_tolerance=1e-8)


def stream(count=12, seed=7, rate=400.0):
    rng = np.random.default_rng(seed)
    keys, spec = [], []
    for _ in range(count):
        if keys and rng.random() < 0.3:
            key = keys[int(rng.integers(len(keys)))]
        else:
            n = int(round(math.exp(rng.uniform(math.log(16), math.log(32)))))
            key = (n, round(float(rng.uniform(100e6, 300e6)) / 1e6, 3), int(rng.integers(1 << 30)))
            keys.append(key)
        ports = [int(p) for p in rng.choice(21, size=int(rng.integers(1, 4)), replace=False)]
        spec.append((key, ports))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    return keys, spec, arrivals


def inputs(key):
    n, fmhz, seed = key
    eps = W.random_permittivity((n, n, n), seed).reshape(-1)
    centers = W.cube_cells(n, PITCH)
    dip = W.Dipole(center=(-0.1, 0.5 * n * PITCH, 0.5 * n * PITCH))
    k0 = W.k0_of(fmhz * 1e6)
    C = dip.impedance_matrix(k0)
    B, D = W.coupling_blocks(dip, centers, PITCH, eps, k0)
    return eps, centers, dip, k0, C, B, D


def make_system(data):
    eps, centers, dip, k0, C, B, D = data
    s = vie.System(centers, [vie.MaterialSpec(1, 1.5)], dip.segments, k0, pitch=PITCH)
    s.set_permittivities(eps)
    s.set_blocks(C=C, B=B, D=D)
    return s


def rhs_for(key, port):
    b = np.zeros(21 + 3 * key[0] ** 3, complex)
    b[port] = 1.0
    return b


@pytest.fixture(scope="module")
def workload():
    keys, spec, arrivals = stream()
    data = {k: inputs(k) for k in keys}
    return keys, spec, arrivals, data


def requests(workload):
    keys, spec, arrivals, data = workload
    return [serve.Request(rid, float(t), key, (key[0],) * 3, 21, (lambda d=data[key]: make_system(d)),
                          [rhs_for(key, p) for p in ports])
            for rid, ((key, ports), t) in enumerate(zip(spec, arrivals))]


@pytest.fixture(scope="module")
def alone(workload):
    """Each distinct (key, port) solved by itself: a fresh system, one job."""
    keys, spec, arrivals, data = workload
    out = {}
    for key in keys:
        s = serve.ready_system(make_system(data[key]))  # as the server readies its systems
        for p in sorted({p for k, ports in spec if k == key for p in ports}):
            batch = vie.Batch()
            jid = batch.submit(s, rhs_for(key, p), precision=vie.MIXED, **OPTS)
            batch.wait(jid)
            out[(key, p)] = batch.result(jid)
            batch.close()
        s.close()
    return out


def check_results(srv, spec, alone):
    for req in srv.done:
        ports = spec[req.rid][1]
        for p, (ok, x, rep, _) in zip(ports, req.results):
            ok1, x1, rep1, _ = alone[(req.key, p)]
            err = float(np.max(np.abs(x - x1)) / np.max(np.abs(x1)))
            print(f"request {req.rid} {req.key} port {p}: {rep.iterations} it (alone {rep1.iterations}), "
                  f"in the mix vs alone {err:.1e}", flush=True)
            assert ok and ok1 and rep.relative_residual <= 1e-8
            assert rep.iterations == rep1.iterations
            assert err <= 1e-12


def test_served_results_equal_solo_solves_and_meet_the_reference(workload, alone, ref):
    keys, spec, arrivals, data = workload
    srv = serve.Server(8 * 10 ** 9, precision=vie.MIXED, share=False, **OPTS)
    try:
        rep = srv.run(requests(workload))
        print("serve report", {k: rep[k] for k in ("requests", "solves", "cache_hits", "system_builds",
                                                   "p50_ms", "p99_ms", "converged")}, flush=True)
        assert rep["requests"] == len(spec) and rep["rejected"] == 0
        assert rep["converged"] == sum(len(p) for _, p in spec)
        assert rep["system_builds"] == len(keys) and rep["cache_hits"] == len(spec) - len(keys)
        from nec2c_amd import timing
        from schema_check import check
        doc = srv.timing_document()
        assert check(doc, timing.schema()) == []
        assert {"fft_build", "fft_upload"} <= {p["name"] for p in doc["phases"]}
        assert doc["serving"]["solves"] == rep["solves"]
        done = sorted(srv.done, key=lambda r: r.admitted)
        assert any(b.admitted < a.finished for a, b in zip(done, done[1:])), "no request overlapped another"
        check_results(srv, spec, alone)
        # every result solves the reference's operator to the tolerance
        for req in srv.done:
            eps, centers, dip, k0, C, B, D = data[req.key]
            n = req.key[0]
            xyz = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
            op = ref.plan((n, n, n), xyz, PITCH, k0, eps[0])
            op.set_permittivities(eps)
            rc = op.coupled(C, B, D)
            for p, (ok, x, _, _) in zip(spec[req.rid][1], req.results):
                b = rhs_for(req.key, p)
                r = float(np.linalg.norm(b - rc.apply(x)) / np.linalg.norm(b))
                assert r <= 1.05e-8, (req.rid, p, r)
    finally:
        srv.close()


def test_served_under_a_one_model_budget(workload, alone):
    keys, spec, arrivals, data = workload
    big = max(keys, key=lambda k: k[0])
    one = serve.system_device_bytes((big[0],) * 3, 21) + 3 * serve.job_device_bytes((big[0],) * 3, 21,
                                                                                   OPTS["restart"])
    srv = serve.Server(one, precision=vie.MIXED, share=False, **OPTS)
    peak = []
    step = srv.step
    srv.step = lambda arrived: (step(arrived), peak.append(srv.used))[0]
    try:
        rep = srv.run(requests(workload))
        assert rep["requests"] == len(spec) and rep["rejected"] == 0
        assert max(peak) <= srv.budget
        check_results(srv, spec, alone)
    finally:
        srv.close()


def test_served_with_shared_subspaces(workload, alone, ref):
    keys, spec, arrivals, data = workload
    srv = serve.Server(8 * 10 ** 9, precision=vie.MIXED, **OPTS)
    try:
        rep = srv.run(requests(workload))
        published, adopted = srv.batch.shared_stats()
        print("serve report (shared)", {k: rep[k] for k in ("requests", "solves", "iterations_total", "converged")},
              "published", published, "adopted", adopted, flush=True)
        assert rep["requests"] == len(spec) and rep["converged"] == sum(len(p) for _, p in spec)
        assert published >= 1 and adopted >= 1
        steps_mix = steps_alone = 0
        for req in srv.done:
            eps, centers, dip, k0, C, B, D = data[req.key]
            n = req.key[0]
            xyz = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
            op = ref.plan((n, n, n), xyz, PITCH, k0, eps[0])
            op.set_permittivities(eps)
            rc = op.coupled(C, B, D)
            for p, (ok, x, r_, _) in zip(spec[req.rid][1], req.results):
                ok1, x1, rep1, _ = alone[(req.key, p)]
                zerr = abs(1 / x[p] - 1 / x1[p]) / abs(1 / x1[p])
                b = rhs_for(req.key, p)
                r = float(np.linalg.norm(b - rc.apply(x)) / np.linalg.norm(b))
                print(f"request {req.rid} port {p}: {r_.iterations} it (alone {rep1.iterations}), impedance vs alone "
                      f"{zerr:.1e}, reference residual {r:.2e}", flush=True)
                assert ok and zerr <= 1e-6 and r <= 1.05e-8
                steps_mix += r_.iterations
                steps_alone += rep1.iterations
        assert steps_mix < steps_alone
    finally:
        srv.close()
