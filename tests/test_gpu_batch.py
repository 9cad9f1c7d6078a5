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


@pytest.fixture(autouse=True, scope="module")
def separate_passes():
    """Solo solves may tune to the fused z+y passes, and batches (multi-job
    launches) run their own launch shape; the bit-for-bit comparisons here
    are between the multi-job and the single-job kernels of the same passes
    and shapes, so these systems tune without fusion and batches keep the
    solo-tuned shapes."""
    import os
    keys = {"VIE_HIP_FUSE": "0", "VIE_HIP_MJ_TUNED": "1"}
    old = {k: os.environ.get(k) for k in keys}
    os.environ.update(keys)
    yield
    for k, v in old.items():
        if v is None:
            del os.environ[k]
        else:
            os.environ[k] = v

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


def test_shared_subspaces_across_ports(ref):
    """vie_hip_batch_share_subspaces: the seed port of each model publishes the
    recycled subspace of its first cycle, the model's other ports adopt it and
    start with a recycled cycle.  Every job converges, meets 1e-8 on the
    reference operator and agrees with its solo (unshared) solve within the
    solver tolerance; the adopting jobs need fewer Arnoldi steps than alone.
    After the operator changes, the old subspace is not adopted."""
    cs = [c for c in W.batch_cases(256, seed=4, kh=0.1) if len(c.ports) >= 3 and np.prod(c.shape) <= 24 ** 3][:5]
    opts = dict(restart=120, maximum_iterations=3000, relative_tolerance=1e-8)
    built = [build(c) for c in cs]
    try:
        for prec, rk in ((vie.MIXED, 0), (vie.FP64, 20)):
            batch = vie.Batch(share_subspaces=True)
            try:
                seeds = [batch.submit(s, full_rhs(case, s, 0), precision=prec, recycle=rk, **opts)
                         for case, (s, _, _) in zip(cs, built)]
                batch.wait_all()
                seed_res = [batch.result(j) for j in seeds]
                rest = [(ci, port, batch.submit(s, full_rhs(case, s, port), precision=prec, recycle=rk, **opts))
                        for ci, (case, (s, _, _)) in enumerate(zip(cs, built)) for port in range(1, len(case.ports))]
                batch.wait_all()
                rest_res = {jid: batch.result(jid) for *_, jid in rest}
                published, adopted = batch.shared_stats()
            finally:
                batch.close()
            print(f"precision {prec}: published {published}, adopted {adopted} of {len(rest)}", flush=True)
            assert published == len(cs) and adopted == len(rest)
            assert all(r[0] for r in seed_res)
            steps_shared = steps_solo = 0
            for ci, port, jid in rest:
                case, (s, blocks, eps) = cs[ci], built[ci]
                ok, x, rep, _ = rest_res[jid]
                b = full_rhs(case, s, port)
                ok1, x1, rep1 = s.solve(b, precision=prec, recycle=rk, **opts)
                fi = case.ports[port]
                zerr = abs(1 / x[fi] - 1 / x1[fi]) / abs(1 / x1[fi])
                r = ref_residual(ref, case, blocks, eps, b, x)
                print(f"model {ci} {case.shape} port {port}: shared {rep.iterations} steps, alone {rep1.iterations}; "
                      f"reference residual {r:.2e}, impedance vs alone {zerr:.1e}", flush=True)
                assert ok and ok1 and r <= 1.05e-8 and zerr <= 1e-6
                steps_shared += rep.iterations
                steps_solo += rep1.iterations
            assert steps_shared < 0.9 * steps_solo
        # a changed operator (new permittivity) does not adopt the old subspace
        case, (s, blocks, eps) = cs[0], built[0]
        eps2 = np.full_like(eps, 0.8 * case.eps_r + 0j)
        s.set_permittivities(eps2)
        batch = vie.Batch(share_subspaces=True)
        try:
            jid = batch.submit(s, full_rhs(case, s, 1), precision=vie.MIXED, **opts)
            batch.wait_all()
            ok, x, rep, _ = batch.result(jid)
            assert batch.shared_stats() == (1, 0)
        finally:
            batch.close()
        assert ok and ref_residual(ref, case, blocks, eps2, full_rhs(case, s, 1), x) <= 1.05e-8
    finally:
        for s, _, _ in built:
            s.close()
