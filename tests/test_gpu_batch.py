"""Batched device solves (vie_hip_batch_*): jobs of uneven grids, several
right-hand sides each, FP64 and mixed precision, all in flight together.

Each result must equal the same solve run alone through dielectric_fft_solve
(same kernels, deterministic reductions: tolerance 1e-12 relative, and the
iteration counts must match), and must satisfy the 1e-8 relative residual
of the reference's own coupled operator (oracle/_ref) on the same system.
The batch advances the matvec steps of its running cycles as multi-job steps
(one launch per kernel for every member), so the same comparison covers the
multi-job kernels against the single-job ones."""
import numpy as np
import pytest

from nec2c_amd import vie, workloads as W

pytestmark = pytest.mark.gpu

OPTS = dict(restart=30, maximum_iterations=600, relative_tolerance=1e-8)


def build(case):
    centers = case.centers()
    eps = np.full(len(centers), case.eps_r, dtype=np.complex128)
    s = vie.System(centers, [vie.MaterialSpec(1, case.eps_r)], case.dipole.segments, case.k0, pitch=case.pitch)
    C = case.dipole.impedance_matrix(case.k0)
    B, D = W.coupling_blocks(case.dipole, centers, case.pitch, eps, case.k0)
    s.set_blocks(C=C, B=B, D=D)
    return s, (C, B, D), eps


def full_rhs(case, s, port):
    b = np.zeros(s.order, dtype=np.complex128)
    b[:case.dipole.segments] = case.rhs(port)
    return b


def ref_residual(ref, case, blocks, eps, b, x):
    xyz = np.stack(np.meshgrid(*[np.arange(v) for v in case.shape], indexing="ij"), -1).reshape(-1, 3)
    p = ref.plan(case.shape, xyz, case.pitch, case.k0, eps[0])
    op = p.coupled(*blocks)
    r = b - op.apply(x)
    return float(np.linalg.norm(r) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def cases():
    cs = W.batch_cases(6, seed=11, lo=4, hi=10, max_ports=3, seg_lo=7, seg_hi=13)
    built = [build(c) for c in cs]
    yield cs, built
    for s, _, _ in built:
        s.close()


def test_batch_matches_solo_and_reference(cases, ref):
    cs, built = cases
    jobs = []
    batch = vie.Batch()
    try:
        # all right-hand sides of all systems in flight, alternating precision
        for ci, (case, (s, blocks, eps)) in enumerate(zip(cs, built)):
            for port in range(len(case.ports)):
                prec = vie.MIXED if (ci + port) % 2 else vie.FP64
                b = full_rhs(case, s, port)
                jobs.append((ci, port, prec, b, batch.submit(s, b, precision=prec, **OPTS)))
        print("jobs in flight", batch.running, flush=True)
        batch.wait_all()
        results = {jid: batch.result(jid) for *_, jid in jobs}
    finally:
        batch.close()
    for ci, port, prec, b, jid in jobs:
        case, (s, blocks, eps) = cs[ci], built[ci]
        ok, x, rep, sec = results[jid]
        ok1, x1, rep1 = s.solve(b, precision=prec, **OPTS)
        print(f"case {ci} {case.shape} port {port} prec {prec}: it {rep.iterations} solo {rep1.iterations} "
              f"rel {rep.relative_residual:.2e} t {sec * 1e3:.1f} ms", flush=True)
        assert ok and ok1
        assert rep.iterations == rep1.iterations
        assert np.max(np.abs(x - x1)) <= 1e-12 * np.max(np.abs(x1))
        assert ref_residual(ref, case, blocks, eps, b, x) <= 1.05e-8


def test_batch_reuses_workspaces_and_streams_results(cases):
    cs, built = cases
    s, _, _ = built[0]
    case = cs[0]
    batch = vie.Batch()
    try:
        batch.reserve(2, s, **OPTS)
        b = full_rhs(case, s, 0)
        first = batch.submit(s, b, **OPTS)
        seen = []
        while len(seen) < 1:
            ready, running = batch.progress()
            seen += ready
        assert seen == [first]
        ok, x, rep, _ = batch.result(first)
        # a second wave through the same pooled workspace gives the same answer
        again = batch.submit(s, b, **OPTS)
        batch.wait(again)
        ok2, x2, rep2, _ = batch.result(again)
        assert ok and ok2 and rep.iterations == rep2.iterations
        assert np.array_equal(x, x2)
    finally:
        batch.close()


def test_batch_rejects_bad_input(cases):
    cs, built = cases
    s, _, _ = built[0]
    batch = vie.Batch()
    try:
        with pytest.raises(RuntimeError):
            batch.submit(s, np.ones(s.order), restart=0)
        with pytest.raises(RuntimeError):
            batch.submit(s, np.ones(s.order), restart=vie.KMAX_RESTART + 1, maximum_iterations=1000)
    finally:
        batch.close()


def test_batch_at_configuration_size(ref):
    """The first 16 models of the batched workload (seed 4: grids 8-48 per
    axis, cubes and cuboids, eps_r 2-12, 50-500 MHz, dipoles of 11-31
    segments, 1-8 ports), every port in flight at once with restart 120:
    each job equals its solo solve and meets 1e-8 on the reference operator."""
    cs = W.batch_cases(256, seed=4, kh=0.1)[:16]
    opts = dict(restart=120, maximum_iterations=3000, relative_tolerance=1e-8)
    built = [build(c) for c in cs]
    batch = vie.Batch()
    jobs = []
    try:
        for ci, (case, (s, blocks, eps)) in enumerate(zip(cs, built)):
            for port in range(len(case.ports)):
                prec = vie.MIXED if (ci + port) % 2 else vie.FP64
                b = full_rhs(case, s, port)
                jobs.append((ci, port, prec, b, batch.submit(s, b, precision=prec, **opts)))
        print(f"{len(jobs)} jobs of {len(cs)} models in flight: {batch.running}", flush=True)
        batch.wait_all()
        results = {jid: batch.result(jid) for *_, jid in jobs}
        # the matvec steps ran as multi-job steps: one launch per kernel for many jobs
        group_steps, job_steps = batch.group_stats()
        print(f"multi-job steps {group_steps} covering {job_steps} job steps", flush=True)
        assert group_steps > 0 and job_steps >= 4 * group_steps
    finally:
        batch.close()
    ops = {}
    for ci, port, prec, b, jid in jobs:
        case, (s, blocks, eps) = cs[ci], built[ci]
        ok, x, rep, sec = results[jid]
        ok1, x1, rep1 = s.solve(b, precision=prec, **opts)
        if ci not in ops:
            xyz = np.stack(np.meshgrid(*[np.arange(v) for v in case.shape], indexing="ij"), -1).reshape(-1, 3)
            plan = ref.plan(case.shape, xyz, case.pitch, case.k0, eps[0])
            ops[ci] = (plan, plan.coupled(*blocks))
        r = float(np.linalg.norm(b - ops[ci][1].apply(x)) / np.linalg.norm(b))
        print(f"model {ci} {case.shape} port {port} prec {prec}: it {rep.iterations} solo {rep1.iterations} "
              f"reference residual {r:.2e}", flush=True)
        assert ok and ok1 and rep.iterations == rep1.iterations
        assert np.max(np.abs(x - x1)) <= 1e-12 * np.max(np.abs(x1))
        assert r <= 1.05e-8
    for s, _, _ in built:
        s.close()
