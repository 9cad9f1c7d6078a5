#!/usr/bin/env python3
"""Benchmark of the MI355X VIE FFT-GMRES.

Default workload, the headline (one "step" = one mixed-precision solve to a
1e-8 true residual from x0 = 0): a 128 x 128 x 128 voxel slab (1 cm voxels,
eps_r 6, 0.01 S/m; 6.29 M volume unknowns, FFT grid padded to 256^3) driven by
a 4-element wire Yagi (84 segments) 0.1 m from its x = 0 face, at 200 MHz, in
free space.  The operator is the reference's coupled [C B; D V]; C/B/D are
synthetic blocks of the reference's shape (nec2c_amd/workloads.py).

Reported: mixed-precision (FP32 cycles, FP64 true residual) solves/s as the
value, FP64 solves/s beside it (the reference's restarted GMRES, and the same
recycled subspace as the mixed solve), matvecs/s, p50/p99 solve latency, the
roofline of the fused matvec (algorithmic bytes of the pruned five-pass model
plus the cross blocks, over HIP-event times of whole applies on the system
stream), the roofline of the dominant Arnoldi kernel, each GPU solution's
true residual on the reference's own operator (oracle/_ref), and the
reference's CPU FFT-GMRES on the same system on a bounded sample.

--workload sweep: the 64^3 random-permittivity block + 21-segment dipole over
the 100-400 MHz sweep (one step = one frequency point); the other workloads
(batch, serve, matvec, farfield, rows, columns, cross_slab, point, mesh) are
listed by --help.

All inputs (spectrum, permittivity, C/B/D, right-hand side) are resident in
HBM before the timed region.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_64cube_matvec.json")


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE, each scaled by its reading on a copy of known
    size); None when that kernel variant was not profiled."""
    try:
        with open(TRAFFIC_FILE) as f:
            table = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    name = kernel.split(" ")[0].replace(" ", "")
    for k, v in table.items():
        if k.replace(" ", "") == name:
            return v["bytes_per_launch"]
    return None


def profiled_traffic_sum(workload, precision):
    """HBM bytes of one fused matvec (the sum over its kernels of FETCH_SIZE
    and WRITE_SIZE per launch, each scaled by its reading on a copy of known
    size) from the committed counter summary profiles/hbm_traffic_<workload>.json;
    None when that workload was not profiled."""
    try:
        with open(os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.json")) as f:
            return json.load(f)["matvec_bytes"][precision]
    except (OSError, ValueError, KeyError):
        return None


METRIC = "VIE matvecs/s & solves/s at 1e-8 residual, mixed vs FP64 GPU; p50/p99 latency"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def sweep_frequencies(count, offset=0):
    full = np.linspace(100e6, 400e6, 101)
    idx = np.linspace(0, 100, count + offset).round().astype(int)
    return [float(full[i]) for i in idx]


def build_case(n, freq, eps, dipole, cells_xyz, centers, pitch, opts):
    from nec2c_amd import vie, workloads as W
    k0 = W.k0_of(freq)
    sys_ = vie.System(centers, [vie.MaterialSpec(1, 1.5)], dipole.segments, k0, pitch=pitch)
    sys_.set_permittivities(eps.reshape(-1))
    C = dipole.impedance_matrix(k0)
    B, D = W.coupling_blocks(dipole, centers, pitch, eps.reshape(-1), k0)
    sys_.set_blocks(C=C, B=B, D=D)
    b = np.zeros(sys_.order, dtype=np.complex128)
    b[: dipole.segments] = dipole.rhs()
    sys_.upload_rhs(b)
    sys_.reserve(**opts)  # blocks, permittivity, basis and vectors resident before timing
    return sys_, (C, B, D, b, k0)


def pass_kernel_name(pas, P, v, E=4096, real="float"):
    """Kernel instantiation of volume pass `pas` at launch variant v
    (lines per workgroup and workgroup size as the library's tile rules; the
    x pass takes twice the elements and threads)."""
    cap = (8, 16, 32, 16, 8)[pas]
    wide = 2 if pas == 2 else 1
    l0 = min(wide * E // P, cap)
    lw = (l0, max(l0 // 2, 1), max(l0 // 4, 1), max(l0 // 2, 1))[v & 3]
    wg = wide * (256, 128, 64, 256)[v & 3]
    tf = lambda b: "true" if b else "false"
    if pas == 2 and v >= 16:
        lw, waves = (32, 2) if v & 1 else (16, 2 if v & 2 else 3)
        return (f"k_xpass_r16<{lw},{tf(v & 4)},{tf(v & 8)},{waves}> (register-resident X FFT + 3x3 spectral "
                f"multiply + inverse X)")
    if pas in (1, 3) and v >= 4:
        lw = (16 if real == "float" else 8) * (2 if v == 5 else 1)
        return f"k_ypass_r16<{real},{lw},{tf(pas == 3)}>"
    pair = f",{tf(v & 4)},{tf(v & 8)}" if pas == 2 else ""
    kern = ("k_zfwd<{R},{P},{lw},{wg}>", "k_ypass<{R},{P},{lw},{wg},false>",
            "k_xpass<{R},{P},{lw},{wg}{pair}> (X FFT + 3x3 spectral multiply + inverse X)",
            "k_ypass<{R},{P},{lw},{wg},true>", "k_zinv<{R},{P},{lw},{wg}>")[pas]
    return kern.format(R=real, P=P, lw=lw, wg=wg, pair=pair)


def cpu_baseline(n, pitch, eps, case, iterations_needed, restart, sample_iters=10, x_sols=None, feed=None,
                 gain_args=None):
    """The reference's own code (oracle/_ref): vie_fft_create + vie_gmres on the
    same coupled system for one frequency point, bounded to ``sample_iters``
    GMRES iterations and scaled to a full solve by the iteration count the GPU
    FP64 solve (the same algorithm) needed at that point.  The scaling is
    optimistic for the CPU: later Arnoldi steps orthogonalise against more
    vectors than the sampled first ones.  With ``x_sols`` (GPU solutions of
    that point by precision) the reference also checks them: their true
    residual on its operator, and, after its own vie_gmres solve to 1e-8 (the
    FP64 oracle), their feed impedance, conductor currents and far-field gain
    against its solution (untimed)."""
    from oracle import oracle as O
    if not O.reference_available():
        return None, None
    R = O.ReferenceLibrary()
    cores = R.threads(0)
    C, B, D, b, k0 = case
    xyz = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
    t0 = time.perf_counter()
    plan = R.plan((n, n, n), xyz, pitch, k0, eps.reshape(-1)[0])
    plan.set_permittivities(eps.reshape(-1))
    t_build = time.perf_counter() - t0
    rc = plan.coupled(C, B, D)
    t1 = time.perf_counter()
    y = rc.apply(b)  # one warm matvec (also times one)
    t_mv = time.perf_counter() - t1
    t2 = time.perf_counter()
    ok, x, rep = rc.gmres(b, restart=restart, maxit=sample_iters, rtol=1e-8)
    t_iters = time.perf_counter() - t2
    del y
    per_it = t_iters / max(rep.iterations, 1)
    total = per_it * iterations_needed   # preparation excluded, as in the GPU value
    check = None
    if x_sols:
        bnorm = float(np.linalg.norm(b))
        check = {f"residual_on_reference_operator_{tag}": float(np.linalg.norm(b - rc.apply(x)) / bnorm)
                 for tag, x in x_sols.items()}
        okr, xr, repr_ = rc.gmres(b, restart=restart, maxit=20 * iterations_needed, rtol=1e-8)
        check["reference_solve"] = {"converged": bool(okr), "iterations": int(repr_.iterations),
                                    "relative_residual": float(repr_.relative_residual)}
        c = len(C)
        for tag, x in x_sols.items():
            check[f"feed_impedance_rel_err_{tag}_vs_reference"] = float(abs(1 / x[feed] - 1 / xr[feed])
                                                                        / abs(1 / xr[feed]))
            check[f"conductor_current_rel_err_{tag}_vs_reference"] = float(np.max(np.abs(x[:c] - xr[:c]))
                                                                           / np.max(np.abs(xr[:c])))
            if gain_args is not None:
                check[f"far_field_gain_max_db_{tag}_vs_reference"] = gain_difference_db(x, xr, k0, *gain_args)
        log("reference check: " + ", ".join(f"{k} {v:.2e}" for k, v in check.items() if isinstance(v, float)))
    return {
        "value": 1.0 / total,
        "unit": "solves/s",
        "cores": cores,
        "kind": "reference",
        "sample": (f"1 frequency point ({k0 * 299792458.0 / 2 / math.pi / 1e6:.1f} MHz) of the same {n}^3 coupled "
                   f"system on the reference's FFT operator (builtin FFT, OpenMP): the first "
                   f"{rep.iterations} vie_gmres({restart}) iterations {t_iters:.2f} s, scaled to the "
                   f"{iterations_needed} iterations the same algorithm needs there; like the GPU value it "
                   f"excludes preparation (reference vie_fft_create {t_build:.2f} s, build_s); one matvec "
                   f"{t_mv * 1e3:.1f} ms; "
                   "[C B; D] products restated around the reference volume operator (its prepare caps at 512 MiB)"),
        "matvec_ms": t_mv * 1e3,
        "build_s": t_build,
        "seconds_per_iteration": per_it,
        "iterations_sampled": rep.iterations,
        "iterations_scaled_to": iterations_needed,
    }, check


def gain_difference_db(x_a, x_b, k0, centers, pitch, wire_points, wire_delta, floor_db=40.0, step_deg=5.0):
    """Largest |gain difference| in dB between two solutions of the same
    coupled system over a step_deg theta/phi grid: the total far field of the
    wire (each segment's solved current as a constant current, NEC's
    amperes-per-wavelength coefficient) and the volume current, through
    vie_hip_total_far_field; directions more than floor_db below the pattern
    maximum are left out (nulls)."""
    from nec2c_amd import vie
    wlam = 2 * math.pi / k0
    th, ph = np.meshgrid(np.radians(np.arange(0.0, 180.0 + 1e-9, step_deg)),
                         np.radians(np.arange(0.0, 360.0, step_deg)), indexing="ij")
    th, ph = th.reshape(-1), ph.reshape(-1)
    nseg = len(wire_points)
    seg = (wire_points[:, 0] / wlam, wire_points[:, 1] / wlam, wire_points[:, 2] / wlam,
           np.broadcast_to(np.asarray(wire_delta, dtype=np.float64) / wlam, (nseg,)).copy(), np.zeros(nseg),
           np.zeros(nseg), np.ones(nseg))
    sizes = np.full_like(centers, pitch)
    gains = []
    for x in (x_a, x_b):
        cur = np.zeros(3 * nseg, dtype=np.complex128)
        cur[0::3] = x[:nseg] / wlam
        et, ep = vie.total_far_field(k0, wlam, th, ph, seg, cur, centers=centers, sizes=sizes,
                                     volume_current=x[nseg:])
        gains.append(np.abs(et) ** 2 + np.abs(ep) ** 2)
    ga, gb = gains
    keep = gb > gb.max() * 10 ** (-floor_db / 10)
    return float(np.max(np.abs(10 * np.log10(ga[keep] / gb[keep]))))


def _case_blocks(case):
    """Host-side inputs of one batched model (the caller's data)."""
    from nec2c_amd import workloads as W
    centers = case.centers()
    eps = np.full(len(centers), case.eps_r, dtype=np.complex128)
    C = case.dipole.impedance_matrix(case.k0)
    B, D = W.coupling_blocks(case.dipole, centers, case.pitch, eps, case.k0)
    return centers, C, B, D


def bench_batch(args):
    """256 independent models: grids 8-48 per axis (cubes and cuboids), eps_r
    U[2, 12], 50-500 MHz, a dipole of 11-31 segments with 1-8 feed ports (one
    right-hand side each), all solved to 1e-8, mixed and FP64, with at most
    --inflight jobs on the device at once (continuous batching)."""
    from concurrent.futures import ThreadPoolExecutor
    from nec2c_amd import vie, serve, workloads as W
    cases = W.batch_cases(args.batch_count, seed=4, kh=args.batch_kh)
    opts = dict(restart=args.restart, maximum_iterations=args.max_iterations, relative_tolerance=1e-8)
    log(f"batch: {len(cases)} models, {sum(len(c.ports) for c in cases)} right-hand sides")
    t0 = time.perf_counter()
    with ThreadPoolExecutor(16) as ex:
        inputs = list(ex.map(_case_blocks, cases))
    t_host = time.perf_counter() - t0
    systems, jobs, feeds, job_case = [], [], [], []
    t1 = time.perf_counter()
    for case, (centers, C, B, D) in zip(cases, inputs):
        s = vie.System(centers, [vie.MaterialSpec(1, case.eps_r)], case.dipole.segments, case.k0, pitch=case.pitch)
        s.set_blocks(C=C, B=B, D=D)
        s.reserve(restart=1, maximum_iterations=1, relative_tolerance=1e-8)  # blocks resident, passes tuned
        systems.append(s)
    # port-major order: the first port of every model (its seed), then the
    # second ports, ...: with shared subspaces the later ports of a model
    # start after its seed has published its recycled subspace.  Within a
    # port, the largest models first (longest jobs first: the batch drains
    # with small jobs, not with a few large ones running alone).
    by_size = sorted(range(len(cases)), key=lambda i: -int(np.prod(cases[i].shape)))
    for port in range(max(len(c.ports) for c in cases)):
        for i in by_size:
            case, s = cases[i], systems[i]
            if port < len(case.ports):
                b = np.zeros(s.order, dtype=np.complex128)
                b[: case.dipole.segments] = case.rhs(port)
                jobs.append((s, b))
                feeds.append(case.ports[port])
                job_case.append(i)
    t_setup = time.perf_counter() - t1
    # the CPU baseline's sample of jobs keeps its models' host blocks
    sample = [] if args.no_cpu else sorted(np.random.default_rng(7).choice(len(jobs), size=min(args.batch_cpu_jobs,
                                                                                                  len(jobs)),
                                                                        replace=False).tolist())
    kept = {job_case[j]: inputs[job_case[j]][1:] for j in sample}
    del inputs
    log(f"batch: host inputs {t_host:.1f} s, device setup (spectra, blocks, tuning) {t_setup:.1f} s")
    biggest = max(systems, key=lambda q: q.order)
    out, group_steps, shared = {}, {}, {}
    k_fp64 = args.recycle if args.recycle > 0 else 16
    runs = (("mixed", vie.MIXED, True, args.recycle),          # the batched form: shared recycled subspaces
            ("fp64", vie.FP64, False, 0),                       # the reference's GMRES, one job per port
            ("fp64_shared", vie.FP64, True, k_fp64),            # FP64 with the same shared subspaces
            ("mixed_unshared", vie.MIXED, False, args.recycle))  # mixed, each job on its own
    for name, prec, share, rk in runs:
        b = vie.Batch()
        b.reserve(args.inflight, biggest, **opts)
        res, secs = serve.run_batch(jobs, precision=prec, inflight=args.inflight, batch=b,
                                    inner_reduction=args.inner_reduction, recycle=rk, share=share, **opts)
        group_steps[name] = b.group_stats()
        shared[name] = b.shared_stats()
        b.close()
        out[name] = (res, secs)
        log(f"batch {name}: {len(jobs)} solves in {secs:.2f} s ({sum(r[0] for r in res)} converged, "
            f"{sum(r[2].iterations for r in res)} Arnoldi steps, subspaces published/adopted {shared[name]})")
    rm, tm = out["mixed"]
    rd, td = out["fp64"]
    z_err = max(abs(1 / xm[1][f] - 1 / xd[1][f]) / abs(1 / xd[1][f]) for xm, xd, f in zip(rm, rd, feeds))
    lat_m = np.array([r[3] for r in rm])
    lat_d = np.array([r[3] for r in rd])
    result = {
        "metric": METRIC,
        "value": len(jobs) / tm,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": len(jobs),
        "warmup": 0,
        "ms_per_step": 1e3 * tm / len(jobs),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 inner (fp32 complex) / c128 outer (fp64 complex)",
        "data": "synthetic",
        "config": {
            "workload": f"batched: {len(cases)} models, grids 8-48 per axis (cubes/cuboids), eps_r U[2,12], "
                        f"50-500 MHz, k0 sqrt(eps_r) h = {args.batch_kh:g}, dipoles of 11-31 segments, 1-8 ports each; "
                        f"one GMRES({args.restart}) job per port to 1e-8, {args.inflight} jobs in flight, ports "
                        "submitted port-major with the recycled subspace of each model's first port shared by its "
                        "other ports (mixed); fp64 = the reference's GMRES per port",
            "models": len(cases), "right_hand_sides": len(jobs),
            "cells_total": int(sum(np.prod(c.shape) for c in cases)),
        },
        "fp64_solves_per_s": len(jobs) / td,
        "speedup_mixed_vs_fp64": td / tm,
        "fp64_shared_solves_per_s": len(jobs) / out["fp64_shared"][1],
        "speedup_mixed_vs_fp64_shared": out["fp64_shared"][1] / tm,
        "mixed_unshared_solves_per_s": len(jobs) / out["mixed_unshared"][1],
        "iterations_total": {k: int(sum(r[2].iterations for r in v[0])) for k, v in out.items()},
        "converged": {k: int(sum(r[0] for r in v[0])) for k, v in out.items()},
        "residual_max": {k: max(r[2].relative_residual for r in v[0]) for k, v in out.items()},
        "shared_subspaces_published_adopted": shared,
        "p50_ms_mixed": 1e3 * float(np.percentile(lat_m, 50)),
        "p99_ms_mixed": 1e3 * float(np.percentile(lat_m, 99)),
        "p50_ms_fp64": 1e3 * float(np.percentile(lat_d, 50)),
        "p99_ms_fp64": 1e3 * float(np.percentile(lat_d, 99)),
        "converged_mixed": int(sum(r[0] for r in rm)),
        "converged_fp64": int(sum(r[0] for r in rd)),
        "iterations_mixed_total": int(sum(r[2].iterations for r in rm)),
        "iterations_fp64_total": int(sum(r[2].iterations for r in rd)),
        "matvecs_per_s_mixed": sum(r[2].operator_applications for r in rm) / tm,
        "matvecs_per_s_fp64": sum(r[2].operator_applications for r in rd) / td,
        "residual_max_mixed": max(r[2].relative_residual for r in rm),
        "residual_max_fp64": max(r[2].relative_residual for r in rd),
        "port_impedance_rel_err_mixed_vs_fp64": z_err,
        # multi-job steps: each issued step is one launch per kernel for all
        # of its member jobs (job_steps / group_steps = jobs per launch)
        "multi_job_steps": {k: {"group_steps": g, "job_steps": j, "jobs_per_launch": (j / g if g else None)}
                            for k, (g, j) in group_steps.items()},
        "setup_seconds": {"host_inputs": t_host, "device": t_setup},
        "roofline": None,
        "cpu_baseline": None,
    }
    if sample:
        cpu, check = batch_cpu_baseline(cases, kept, jobs, job_case, feeds, sample, rm, rd, args.restart)
        result["cpu_baseline"], result["oracle_check"] = cpu, check
    print(json.dumps(result), flush=True)
    for s in systems:
        s.close()
    return 0


def batch_cpu_baseline(cases, blocks, jobs, job_case, feeds, sample, res_mixed, res_fp64, restart):
    """The reference's own FFT operator and vie_gmres (oracle/_ref, OpenMP)
    solving a seeded random sample of the batch's jobs to 1e-8, one at a
    time: solves/s over the sample's GMRES time (plan creation excluded, as
    the GPU value excludes its setup).  The same run checks the sampled jobs'
    GPU solutions: their true residual on the reference operator and their
    port impedance against the reference's solution."""
    from oracle import oracle as O
    if not O.reference_available():
        return None, None
    R = O.ReferenceLibrary()
    cores = R.threads(0)
    t_solve, t_plan, iters = 0.0, 0.0, 0
    worst = {"residual_on_reference_operator_mixed": 0.0, "residual_on_reference_operator_fp64": 0.0,
             "port_impedance_rel_err_mixed_vs_reference": 0.0, "port_impedance_rel_err_fp64_vs_reference": 0.0}
    cells = []
    for j in sample:
        case = cases[job_case[j]]
        C, B, D = blocks[job_case[j]]
        b = jobs[j][1]
        xyz = np.stack(np.meshgrid(*[np.arange(v) for v in case.shape], indexing="ij"), -1).reshape(-1, 3)
        t0 = time.perf_counter()
        plan = R.plan(case.shape, xyz, case.pitch, case.k0, complex(case.eps_r))
        op = plan.coupled(C, B, D)
        t1 = time.perf_counter()
        ok, xr, rep = op.gmres(b, restart=restart, maxit=2000, rtol=1e-8)
        t2 = time.perf_counter()
        if not ok:
            log(f"batch cpu baseline: reference solve of job {j} did not converge")
            return None, None
        t_plan += t1 - t0
        t_solve += t2 - t1
        iters += rep.iterations
        cells.append(int(np.prod(case.shape)))
        bn = float(np.linalg.norm(b))
        f = feeds[j]
        for tag, res in (("mixed", res_mixed), ("fp64", res_fp64)):
            x = res[j][1]
            worst[f"residual_on_reference_operator_{tag}"] = max(
                worst[f"residual_on_reference_operator_{tag}"], float(np.linalg.norm(b - op.apply(x)) / bn))
            worst[f"port_impedance_rel_err_{tag}_vs_reference"] = max(
                worst[f"port_impedance_rel_err_{tag}_vs_reference"],
                float(abs(1 / x[f] - 1 / xr[f]) / abs(1 / xr[f])))
        del op, plan
    log("batch reference check: " + ", ".join(f"{k} {v:.2e}" for k, v in worst.items()))
    cpu = {
        "value": len(sample) / t_solve,
        "unit": "solves/s",
        "cores": cores,
        "kind": "reference",
        "sample": (f"{len(sample)} of the {len(jobs)} jobs (seeded uniform draw; models of {min(cells)}-{max(cells)} "
                   f"cells) solved one at a time by the reference's vie_gmres({restart}) on its FFT operator "
                   f"(builtin FFT, OpenMP, FP64) to 1e-8: {iters} iterations in {t_solve:.2f} s of GMRES; plan "
                   f"creation ({t_plan:.2f} s) excluded like the GPU's setup"),
    }
    return cpu, {"jobs_checked": len(sample), **worst}


def bench_serve(args):
    """Poisson arrivals (seeded) of frequency points for models with grids
    16^3-96^3 (log-uniform), random eps_r fields as in the sweep, 1-4 ports, 30%
    of requests repeating an earlier (grid, frequency, eps) key; mixed precision
    to 1e-8 under an HBM budget (288 GB or the device's free memory).  The
    caller-side inputs (eps fields, C/B/D) are generated before the clock
    starts; spectrum builds on cache misses, uploads and solves are served."""
    from concurrent.futures import ThreadPoolExecutor
    from nec2c_amd import vie, serve, workloads as W
    rng = np.random.default_rng(5)
    keys, reqs_spec = [], []
    for _ in range(args.requests):
        if keys and rng.random() < 0.3:
            key = keys[int(rng.integers(len(keys)))]
        else:
            n = int(round(math.exp(rng.uniform(math.log(16), math.log(96)))))
            key = (n, round(float(rng.uniform(100e6, 400e6)) / 1e6, 3), int(rng.integers(1 << 30)))
            keys.append(key)
        reqs_spec.append((key, int(rng.integers(1, 5))))
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, args.requests))
    pitch = 0.5 / 64

    def model_inputs(key):
        n, fmhz, seed = key
        eps = W.random_permittivity((n, n, n), seed).reshape(-1)
        centers = W.cube_cells(n, pitch)
        dip = W.Dipole(center=(-0.1, 0.5 * n * pitch, 0.5 * n * pitch))
        k0 = W.k0_of(fmhz * 1e6)
        C = dip.impedance_matrix(k0)
        B, D = W.coupling_blocks(dip, centers, pitch, eps, k0)
        return key, (eps, centers, dip, k0, C, B, D)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(16) as ex:
        inputs = dict(ex.map(model_inputs, keys))
    log(f"serve: {len(keys)} distinct models, host inputs {time.perf_counter() - t0:.1f} s")

    def factory(key):
        eps, centers, dip, k0, C, B, D = inputs[key]

        def make():
            s = vie.System(centers, [vie.MaterialSpec(1, 1.5)], dip.segments, k0, pitch=pitch)
            s.set_permittivities(eps)
            s.set_blocks(C=C, B=B, D=D)
            return s
        return make

    requests = []
    for rid, ((key, ports), t) in enumerate(zip(reqs_spec, arrivals)):
        dip = inputs[key][2]
        segs = rng.choice(dip.segments, size=ports, replace=False)
        rhs = []
        for seg in segs:
            b = np.zeros(dip.segments + 3 * key[0] ** 3, dtype=np.complex128)
            b[int(seg)] = 1.0
            rhs.append(b)
        requests.append(serve.Request(rid, float(t), key, (key[0],) * 3, dip.segments, factory(key), rhs))
    free, total = vie.memory_info()
    budget = min(288 * 10 ** 9, int(0.9 * free))
    srv = serve.Server(budget, precision=vie.MIXED, restart=args.restart, maximum_iterations=args.max_iterations,
                       inner_reduction=args.inner_reduction, recycle=args.recycle)
    rep = srv.run(requests)
    rep["shared_subspaces_published_adopted"] = srv.batch.shared_stats()
    srv.close()
    result = {
        "metric": METRIC,
        "value": rep["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": rep["solves"],
        "warmup": 0,
        "ms_per_step": 1e3 / rep["solves_per_s"] if rep["solves_per_s"] else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 inner (fp32 complex) / c128 outer (fp64 complex)",
        "data": "synthetic",
        "config": {
            "workload": f"served: {args.requests} Poisson requests ({args.rate:g}/s, seed 5), grids 16^3-96^3 "
                        "log-uniform, random eps_r fields (as the sweep), 100-400 MHz, 1-4 ports, 30% key reuse, "
                        f"mixed GMRES({args.restart}) to 1e-8, HBM budget {budget / 1e9:.0f} GB; each model's "
                        "recycled subspace shared by its ports and later requests (a new model's first port first); "
                        "new systems run the default launch shapes (no per-system tuning)",
            "distinct_models": len(keys),
        },
        "serve": rep,
        "p50_ms": rep["p50_ms"],
        "p99_ms": rep["p99_ms"],
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)
    return 0


SLAB_H = 0.01          # voxel pitch (m) of the slab workload (nec2c_amd/workloads.py yagi_slab)


def slab_geometry(n, ground=False):
    from nec2c_amd import workloads as W
    return W.yagi_slab(n, ground=ground)


def slab_ground_epsilon():
    from nec2c_amd import workloads as W
    er, sigma = W.SLAB_GROUND
    return W.nec_ground_epsilon(er, sigma, W.SLAB_FREQ / 1e6)


def slab_system(n, opts, keep_blocks=False, ground=False):
    """Prepared GPU system of the slab (spectrum, C/B/D, rhs, reserved basis),
    over the real ground when ``ground``.  Returns (system, yagi, k0,
    eps_cell, blocks or None, prepare seconds)."""
    from nec2c_amd import vie, workloads as W
    yagi, k0, eps_cell = slab_geometry(n, ground)
    t0 = time.perf_counter()
    centers = W.cube_cells(n, SLAB_H, origin=W.slab_origin(ground))
    s = vie.System(centers, [vie.MaterialSpec(1, 6.0, 0.0, 0.01)], yagi.segments, k0, pitch=SLAB_H,
                   ground=slab_ground_epsilon() if ground else None)
    t_prep = time.perf_counter() - t0
    log(f"slab: {n}^3 prepared (spectrum{' + reflected spectrum' if ground else ''}) in {t_prep:.1f} s; "
        f"building Yagi blocks (c = {yagi.segments})")
    C = yagi.impedance_matrix(k0)
    B, D = W.coupling_blocks(yagi, centers, SLAB_H, eps_cell, k0, threads=16)
    s.set_blocks(C=C, B=B, D=D)
    b = np.zeros(s.order, dtype=np.complex128)
    b[: yagi.segments] = yagi.rhs()
    s.upload_rhs(b)
    s.reserve(**opts)
    log(f"slab: ready in {time.perf_counter() - t0:.1f} s")
    blocks = (C, B, D, b) if keep_blocks else None
    return s, yagi, k0, eps_cell, blocks, t_prep


def matvec_bytes_bmv(n, P, c, S, ground=False):
    """Algorithmic bytes of one fused coupled matvec by the pruned five-pass
    model: E_vec = 3 (2N + 4 nx ny Pz + 4 nx Py Pz) vector traffic, E_aux =
    3N (x read again for the contrast) + N (permittivity), E_spec = the octant
    spectrum 6 (Px/2+1)(Py/2+1)(Pz/2+1) that the multiply reads (over a real
    ground also the reflected spectrum 6 (Px/2+1)(Py/2+1) Pz), E_cross =
    c^2 + 2 c 3N (C, B and D streamed once); S bytes per complex value."""
    N = n ** 3
    e_vec = 3 * (2 * N + 4 * n * n * P + 4 * n * P * P)
    e_aux = 3 * N + N
    e_spec = 6 * (P // 2 + 1) ** 3 + (6 * (P // 2 + 1) ** 2 * P if ground else 0)
    e_cross = c * c + 2 * c * 3 * N
    return {"volume": (e_vec + e_aux + e_spec) * S, "cross": e_cross * S,
            "total": (e_vec + e_aux + e_spec + e_cross) * S,
            "spectrum_full_6G": 6 * P ** 3 * S}


def slab_ground_table_check(n, k0):
    """The device's reflected-tensor table at the slab's full size against the
    oracle (oracle/ground.py, pinned to the reference's vie_halfspace_cuboids)
    on sampled lateral offsets, every layer sum: the largest difference over
    the table's largest entry."""
    from nec2c_amd import vie, workloads as W
    from oracle import ground as G
    z0 = W.SLAB_GROUND_HEIGHT + 0.5 * SLAB_H
    t0 = time.perf_counter()
    T = vie.ground_table((n, n, n), SLAB_H, z0, k0, slab_ground_epsilon())
    t_dev = time.perf_counter() - t0
    dx = np.array(sorted({0, 1, 2, 7, n // 3, n - 1}))
    dy = np.array(sorted({0, 3, 11, n // 2, n - 1}))
    t1 = time.perf_counter()
    To = G.table((n, n, n), SLAB_H, z0, k0, slab_ground_epsilon(), lateral=(dx, dy))
    t_ora = time.perf_counter() - t1
    Td = T[:, dx][:, :, dy]
    err = float(np.max(np.abs(Td - To)) / np.max(np.abs(T)))
    log(f"ground table: device {t_dev:.2f} s, oracle on {len(dx) * len(dy)} lateral offsets x {2 * n - 1} layer "
        f"sums {t_ora:.1f} s, max difference {err:.2e} of the largest entry")
    return {"max_rel_diff": err, "lateral_offsets": [dx.tolist(), dy.tolist()], "device_table_s": t_dev}


def slab_cpu_leg(n, k0, eps_cell, blocks, x_sols, iterations_needed, restart, sample_iters, ground=False):
    """The reference's own FFT operator (oracle/_ref: vie_fft.c, vie_kernel.c,
    vie_iterative.c built from its sources with OpenMP) on the slab:
    vie_fft_create at 128^3 (timed, reported beside the value), the true
    relative residual of each GPU solution on that operator with the same
    [C B; D] blocks (free space only), and the first ``sample_iters``
    vie_gmres(restart) iterations of the same system, scaled to
    ``iterations_needed``.  Over a real ground the reference has no FFT
    operator (dielectric_ground.c:28-29; its GN 2 path is a dense matrix,
    (6.3 M)^2 entries here): its free-space FFT-GMRES on the same lattice and
    blocks, the fastest path it has at this size, is timed instead."""
    from oracle import oracle as O
    if not O.reference_available():
        return None, {}
    R = O.ReferenceLibrary()
    cores = R.threads(0)
    C, B, D, b = blocks
    xyz = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
    t0 = time.perf_counter()
    plan = R.plan((n, n, n), xyz, SLAB_H, k0, eps_cell)
    t_build = time.perf_counter() - t0
    log(f"cpu: reference vie_fft_create at {n}^3 in {t_build:.1f} s on {cores} threads")
    rc = plan.coupled(C, B, D)
    bnorm = float(np.linalg.norm(b))
    resid = {}
    t_mv = None
    for tag, x in (x_sols.items() if not ground else []):
        t1 = time.perf_counter()
        r = b - rc.apply(x)
        t_mv = time.perf_counter() - t1
        resid[tag] = float(np.linalg.norm(r) / bnorm)
        log(f"cpu: {tag} solution, relative residual on the reference operator {resid[tag]:.3e}")
    if t_mv is None:
        t1 = time.perf_counter()
        rc.apply(b)
        t_mv = time.perf_counter() - t1
    t2 = time.perf_counter()
    ok, x, rep = rc.gmres(b, restart=restart, maxit=sample_iters, rtol=1e-8)
    t_iters = time.perf_counter() - t2
    per_it = t_iters / max(rep.iterations, 1)
    cb = {
        "value": 1.0 / (per_it * iterations_needed),
        "unit": "solves/s",
        "cores": cores,
        "kind": "reference",
        "sample": (("the reference's free-space FFT operator (its FFT path rejects a real ground, "
                    "dielectric_ground.c:28-29) on " if ground else "") +
                   f"the same {n}^3 slab + Yagi system on the reference's FFT operator (builtin FFT, OpenMP, "
                   f"FP64): the first {rep.iterations} vie_gmres({restart}) iterations took {t_iters:.1f} s, "
                   f"scaled to the {iterations_needed} iterations the same FP64 algorithm needs on the GPU "
                   f"(optimistic for the CPU: later steps orthogonalise against more vectors); like the GPU "
                   f"value it excludes preparation (reference vie_fft_create {t_build:.1f} s, reported as "
                   f"build_s); [C B; D] products restated around the reference volume operator, whose own "
                   f"prepare caps storage at 512 MiB"),
        "build_s": t_build,
        "matvec_s": t_mv,
        "seconds_per_iteration": per_it,
        "iterations_sampled": rep.iterations,
        "iterations_scaled_to": iterations_needed,
    }
    return cb, resid


def bench_slab(args):
    """The 128^3 slab (the headline): eps_r 6, 0.01 S/m, 1 cm voxels, 0.1 m
    above a real ground (eps_r 13, 0.005 S/m; --free-space drops it), driven
    by a 4-element Yagi 0.1 m from its x = 0 face, at 200 MHz.  One step is one
    mixed-precision solve of that system to a 1e-8 true residual from x0 = 0.
    FP64 solves (the reference's restarted GMRES, and the same recycled
    subspace as the mixed solve) are timed beside it."""
    from nec2c_amd import vie
    n = args.slab_size
    steps, warmup = args.steps, args.warmup
    opts = dict(restart=args.restart, maximum_iterations=args.max_iterations, relative_tolerance=1e-8)
    ground = not args.free_space
    s, yagi, k0, eps_cell, blocks, t_prep = slab_system(n, opts, keep_blocks=not args.no_cpu, ground=ground)
    c = yagi.segments

    def timed(prec, recycle, count, warm):
        for _ in range(warm):
            s.solve_resident(prec, args.inner_reduction, recycle, **opts)
        vie.synchronize()
        lat, reps = [], []
        t = time.perf_counter()
        for _ in range(count):
            a = time.perf_counter()
            ok, rep = s.solve_resident(prec, args.inner_reduction, recycle, **opts)
            lat.append(time.perf_counter() - a)
            if not ok:
                raise RuntimeError(f"slab solve failed: {vie.STATUS[rep.status]} rel {rep.relative_residual}")
            reps.append(rep)
        vie.synchronize()
        return time.perf_counter() - t, lat, reps

    log(f"slab: {warmup} warm-up + {steps} timed mixed solves")
    s.solve_stats(reset=True)
    t_mixed, lat_mixed, rep_mixed = timed(vie.MIXED, args.recycle, steps, warmup)
    mean_col, _ = s.solve_stats(reset=True)
    rec_k = s.recycled
    x_mixed = s.download_solution()
    log(f"slab mixed: {rep_mixed[0].iterations} iterations, {t_mixed / steps:.3f} s per solve")
    fsteps = max(1, min(args.fp64_steps, steps))
    t_fp64, lat_fp64, rep_fp64 = timed(vie.FP64, 0, fsteps, 1)
    x_fp64 = s.download_solution()
    log(f"slab fp64: {rep_fp64[0].iterations} iterations, {t_fp64 / fsteps:.3f} s per solve")
    k_r = args.recycle if args.recycle > 0 else 16
    t_fp64r, _, rep_fp64r = timed(vie.FP64, k_r, fsteps, 1)
    log(f"slab fp64 recycled: {rep_fp64r[0].iterations} iterations, {t_fp64r / fsteps:.3f} s per solve")

    # fused matvec: HIP events around whole applies on the system's stream
    variants = s.tune()
    ms32, p32 = s.time_apply(vie.MIXED, args.matvec_repeats)
    ms64, p64 = s.time_apply(vie.FP64, args.matvec_repeats)
    lib32, pb32 = s.apply_bytes(vie.MIXED)
    P = 2 * n
    bmv32 = matvec_bytes_bmv(n, P, c, 8, ground)
    bmv64 = matvec_bytes_bmv(n, P, c, 16, ground)
    names = ["zfwd", "yfwd", "xpass", "yinv", "zinv", "coupling"]
    gbs = bmv32["total"] / (ms32 * 1e-3) / 1e9
    # the solve's streaming Arnoldi kernels at the mean basis column of the mixed solves
    jbar = max(1, int(round(mean_col)))
    arn = s.time_arnoldi(vie.MIXED, jbar, args.matvec_repeats)
    arn["gbs_reduce"] = arn["bytes_reduce"] / (arn["ms_reduce"] * 1e-3) / 1e9
    arn["gbs_update"] = arn["bytes_update"] / (arn["ms_update"] * 1e-3) / 1e9

    fi = yagi.feed_index()
    zm, zd = 1 / x_mixed[fi], 1 / x_fp64[fi]
    i_err = float(np.max(np.abs(x_mixed[:c] - x_fp64[:c])) / np.max(np.abs(x_fp64[:c])))
    from nec2c_amd import workloads as W
    gain_db = gain_difference_db(x_mixed, x_fp64, k0, W.cube_cells(n, SLAB_H, origin=W.slab_origin(ground)), SLAB_H,
                                 yagi.points(), yagi.deltas())
    table_check = slab_ground_table_check(n, k0) if ground else None
    log(f"slab far-field gain, mixed vs FP64: max {gain_db:.2e} dB over a 5-degree grid")
    traffic = profiled_traffic_sum(f"{n}slab{'_ground' if ground else ''}_matvec", "fp32")
    result = {
        "metric": METRIC,
        "value": steps / t_mixed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": 1e3 * t_mixed / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 inner (fp32 complex) / c128 outer (fp64 complex)",
        "data": "synthetic",
        "config": {"workload": (f"{n}^3 voxel slab (eps_r 6, 0.01 S/m, 1 cm voxels, {s.order} unknowns, FFT grid "
                                f"padded to {P}^3) + 4-element 84-segment wire Yagi 0.1 m away, 200 MHz, "
                                f"GMRES({args.restart}) to 1e-8; " +
                                (f"0.1 m above a real ground (eps_r 13, 0.005 S/m: eps {slab_ground_epsilon():.4f} "
                                 f"by NEC's GN rule) through the reflected half-space tensor "
                                 f"(vie_halfspace_cuboids, the reference's GN 2 treatment) as a second spectrum; "
                                 f"synthetic free-space wire blocks (NEC's Sommerfeld wire-ground terms are the "
                                 f"caller's)" if ground else "in free space")),
                   "grid": [n, n, n], "padded": [P, P, P], "unknowns": s.order, "conductors": c,
                   "ground": [W.SLAB_GROUND[0], W.SLAB_GROUND[1]] if ground else None},
        "ground_table_check": table_check,
        "fp64_solves_per_s": fsteps / t_fp64,
        "speedup_mixed_vs_fp64": (t_fp64 / fsteps) / (t_mixed / steps),
        "fp64_recycled_solves_per_s": fsteps / t_fp64r,
        "speedup_mixed_vs_fp64_recycled": (t_fp64r / fsteps) / (t_mixed / steps),
        "recycled_subspace_mixed": rec_k,
        "matvecs_per_s_mixed": 1e3 / ms32,
        "matvecs_per_s_fp64": 1e3 / ms64,
        "solve_matvecs_per_s_mixed": sum(r.operator_applications for r in rep_mixed) / t_mixed,
        "p50_ms_mixed": 1e3 * float(np.percentile(lat_mixed, 50)),
        "p99_ms_mixed": 1e3 * float(np.percentile(lat_mixed, 99)),
        "p50_ms_fp64": 1e3 * float(np.percentile(lat_fp64, 50)),
        "p99_ms_fp64": 1e3 * float(np.percentile(lat_fp64, 99)),
        "iterations_mixed": rep_mixed[0].iterations,
        "iterations_fp64": rep_fp64[0].iterations,
        "iterations_fp64_recycled": rep_fp64r[0].iterations,
        "residual_mixed": max(r.relative_residual for r in rep_mixed),
        "residual_fp64": max(r.relative_residual for r in rep_fp64),
        "feed_impedance_ohm_fp64": str(complex(np.round(zd, 6))),
        "feed_impedance_rel_err_mixed_vs_fp64": float(abs(zm - zd) / abs(zd)),
        "conductor_current_rel_err_mixed_vs_fp64": i_err,
        "far_field_gain_max_db_mixed_vs_fp64": gain_db,
        "matvec": {"ms_mixed": ms32, "ms_fp64": ms64,
                   "bytes_bmv_mixed": bmv32, "bytes_bmv_fp64": bmv64,
                   "gbs_bmv_fp64": bmv64["total"] / (ms64 * 1e-3) / 1e9,
                   "pass_ms_mixed": dict(zip(names, p32)), "pass_ms_fp64": dict(zip(names, p64)),
                   "pass_bytes_mixed": dict(zip(names, pb32)),
                   "pass_gbs_mixed": {k: b / (t * 1e-3) / 1e9 for k, b, t in zip(names, pb32, p32) if t > 0},
                   "tuned_variants": variants},
        "setup_seconds": {"prepare_with_spectrum": t_prep},
        "roofline": {
            "bound": "hbm",
            "kernel": ("fused coupled matvec, FP32: five volume passes (k_zfwd, k_ypass, k_xpass, k_ypass "
                       "inverse, k_zinv) + conductor coupling (k_dots, k_couple_out), HIP events around "
                       f"{args.matvec_repeats} whole applies on the system stream"),
            "achieved": gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": gbs / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_launch": bmv32["total"],
        },
        "roofline_solve_kernel": {
            "bound": "hbm",
            "kernel": f"k_dots_dc<float> (delayed-CGS2 reduction at basis column {jbar}, the mean of the mixed "
                      f"solves; ({jbar} + 2) x {s.order} complex float per launch)",
            "achieved": arn["gbs_reduce"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": arn["gbs_reduce"] / HBM_PEAK_GBS, "update_gbs": arn["gbs_update"],
            "ms_reduce": arn["ms_reduce"], "ms_update": arn["ms_update"],
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        log("CPU baseline: the reference's FFT operator at the same size")
        cb, resid = slab_cpu_leg(n, k0, eps_cell, blocks, {"mixed": x_mixed, "fp64": x_fp64},
                                 rep_fp64[0].iterations, args.restart, args.cpu_iterations, ground=ground)
        result["cpu_baseline"] = cb
        if not ground:
            result["residual_on_reference_operator"] = resid
    if ground:
        # the FP64 oracle over a real ground: the reference's GN 2 path is a
        # dense matrix ((6.3 M)^2 entries at this size), so at full size the
        # mixed solution is stated against this FP64 solve; the GN 2 operator
        # and its solves are checked against the reference's dense
        # vie_cuboid_n + vie_halfspace_cuboids on small lattices, and the
        # reflected table here against the oracle (ground_table_check)
        result["oracle_check"] = {
            "fp64_reference_at_full_size": "unavailable over a real ground (dense GN 2 path only)",
            "feed_impedance_rel_err_mixed_vs_fp64": result["feed_impedance_rel_err_mixed_vs_fp64"],
            "small_lattice_parity": "tests/test_gpu_ground.py: operator and solves against the reference's dense "
                                    "product on 6x5x4 and 9x7x5 lattices",
            "free_space_full_size": "bench.py --free-space: both solutions' residual on the reference operator"}
    print(json.dumps(result), flush=True)
    s.close()
    return 0


def bench_matvec(args):
    """Fused operator applies (FP32 and FP64, resident input, Jacobi fused) of
    the first timed sweep point (--geometry sweep) or of the slab (--geometry
    slab), and a plain device copy of known size, the calibration kernel for
    HBM byte counters.  Short by design: counter runs replay every kernel."""
    from nec2c_amd import vie, workloads as W
    opts = dict(restart=4, maximum_iterations=4, relative_tolerance=1e-8)
    if args.geometry == "slab":
        n = args.slab_size
        s, _, _, _, _, _ = slab_system(n, opts, ground=not args.free_space)
        what = f"{n}^3 slab + 84-segment Yagi, 200 MHz, {'free space' if args.free_space else 'over the real ground'}"
    else:
        n = args.size
        pitch = 0.5 / n
        eps = W.random_permittivity((n, n, n), seed=2)
        centers = W.cube_cells(n, pitch)
        dipole = W.Dipole(center=(-0.1, 0.25, 0.25))
        f = sweep_frequencies(args.steps + args.warmup)[args.warmup]
        s, _ = build_case(n, f, eps, dipole, None, centers, pitch, opts)
        what = f"{n}^3 sweep point {f / 1e6:.1f} MHz"
    variants = s.tune()
    out = {"workload": f"{what}, fused apply (Jacobi) x {args.matvec_repeats}", "tuned_variants": variants}
    names = ["zfwd", "yfwd", "xpass", "yinv", "zinv", "coupling"]
    for prec, tag in ((vie.MIXED, "fp32"), (vie.FP64, "fp64")):
        ms, passes = s.time_apply(prec, args.matvec_repeats)
        tot, pb = s.apply_bytes(prec)
        out[tag] = {"ms": ms, "bytes": tot, "pass_ms": dict(zip(names, passes)), "pass_bytes": dict(zip(names, pb)),
                    "pass_gbs": {k: (b / (t * 1e-3) / 1e9 if t > 0 else None) for k, b, t in zip(names, pb, passes)}}
    if args.pass_sweep:
        # every valid launch variant of the y and x passes, the others as tuned
        out["pass_variants"] = {}
        for prec, tag in ((vie.MIXED, "fp32"), (vie.FP64, "fp64")):
            out["pass_variants"][tag] = {}
            for pas in (1, 2, 3):
                rows = {}
                for v in range(32):
                    f64, f32 = list(variants["fp64"]), list(variants["fp32"])
                    (f32 if prec == vie.MIXED else f64)[pas] = v
                    if not s.set_variants(f64, f32):
                        continue
                    _, passes = s.time_apply(prec, args.matvec_repeats)
                    rows[v] = {"kernel": pass_kernel_name(pas, 2 * n, v, 4096 if prec == vie.MIXED else 2048,
                                                          "float" if prec == vie.MIXED else "double").split(" ")[0],
                               "ms": passes[pas]}
                out["pass_variants"][tag][names[pas]] = rows
        s.set_variants(variants["fp64"], variants["fp32"])
    if args.arnoldi_column > 0:
        s.reserve(restart=min(args.arnoldi_column + 1, vie.KMAX_RESTART), maximum_iterations=1000,
                  relative_tolerance=1e-8)
        out["arnoldi"] = {tag: s.time_arnoldi(prec, args.arnoldi_column, args.matvec_repeats)
                          for prec, tag in ((vie.MIXED, "fp32"), (vie.FP64, "fp64"))}
        out["arnoldi"]["j"] = args.arnoldi_column
    cms = vie.copy_probe(args.copy_bytes, 5)
    out["copy"] = {"bytes": args.copy_bytes, "ms": cms, "gbs": 2 * args.copy_bytes / (cms * 1e-3) / 1e9}
    print(json.dumps(out), flush=True)
    s.close()
    return 0


def bench_farfield(args):
    """Scattered far field of a volume current on the n^3 cube of the sweep
    (uniform lattice, edge 0.5 m): the reference's 974-direction scattered
    power and a 1-degree pattern (181 x 360 directions), free space and over
    a PEC plane 0.05 m below the cube.  Synthetic current.  CPU baseline: the
    oracle's numpy restatement (one core) on a bounded sample of directions,
    scaled per cell-direction pair."""
    from nec2c_amd import vie, workloads as W
    n = args.farfield_size
    pitch = 0.5 / n
    centers = W.cube_cells(n, pitch) + np.array([0.0, 0.0, 0.05])
    sizes = np.full_like(centers, pitch)
    rng = np.random.default_rng(11)
    J = rng.standard_normal(3 * len(centers)) + 1j * rng.standard_normal(3 * len(centers))
    k0 = W.k0_of(300e6)
    th, ph = np.meshgrid(np.radians(np.arange(181.0)), np.radians(np.arange(360.0)), indexing="ij")
    pattern = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1).reshape(-1, 3)
    vie.scattered_power(centers, sizes, J, k0)  # warm-up (device buffers, code objects)
    res = {}
    for ground, tag in ((0, "free"), (1, "pec")):
        t = time.perf_counter()
        p = vie.scattered_power(centers, sizes, J, k0, ground=ground)
        t_pow = time.perf_counter() - t
        dirs = pattern if ground == 0 else pattern[pattern[:, 2] >= 0]
        t = time.perf_counter()
        vie.radiation_amplitudes(centers, sizes, J, k0, dirs, ground=ground)
        t_pat = time.perf_counter() - t
        res[tag] = {"power_w": p, "power_ms": t_pow * 1e3, "pattern_directions": len(dirs), "pattern_ms": t_pat * 1e3,
                    "pairs_per_s": len(dirs) * len(centers) / t_pat}
        log(f"farfield {tag}: power {p:.6e} W in {t_pow * 1e3:.2f} ms; {len(dirs)}-direction pattern in "
            f"{t_pat * 1e3:.1f} ms ({res[tag]['pairs_per_s']:.3e} cell-direction pairs/s)")
    cpu = None
    if not args.no_cpu:
        from oracle import farfield as F
        sample = pattern[:: max(1, len(pattern) // 16)][:16]
        t = time.perf_counter()
        F.radiation_amplitudes(centers, sizes, J, k0, sample)
        t_cpu = time.perf_counter() - t
        rate = len(sample) * len(centers) / t_cpu
        cpu = {"value": rate / len(centers), "unit": "directions/s", "cores": 1, "kind": "port",
               "sample": f"{len(sample)} pattern directions x {len(centers)} cells through oracle/farfield.py "
                         f"(numpy, one core, free space) in {t_cpu:.2f} s"}
        log(f"farfield CPU port: {rate:.3e} pairs/s")
    free = res["free"]
    print(json.dumps({
        "metric": "far-field directions/s (radiation amplitude of a volume current)",
        "value": free["pattern_directions"] / (free["pattern_ms"] * 1e-3), "unit": "directions/s",
        "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": free["pattern_ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c128 (fp64 complex)", "data": "synthetic",
        "config": {"workload": f"{n}^3 uniform cube (edge 0.5 m, 0.05 m above z = 0), 300 MHz, random current; "
                               "1-degree pattern (181 x 360) and the 974-direction scattered power",
                   "cells": len(centers)},
        "farfield": res, "cpu_baseline": cpu, "roofline": None,
    }))
    return 0


def bench_rows(args):
    """Volume-to-wire rows of the sweep's coupled system: the n^3 cube (edge
    0.5 m) and the 21-segment dipole 0.1 m from one face, every (segment,
    cell) pair by the reference's 10^3-point rule, free space and over a PEC
    plane 0.05 m below the cube.  Kernel time with the geometry resident
    (HIP events), plus the whole call with the rows copied to the host.  CPU
    baseline: the C restatement oracle/conductor_rows.c with OpenMP over the
    tasks, as the reference runs them, on a bounded sample of cells."""
    from nec2c_amd import vie, workloads as W
    n = args.rows_size
    pitch = 0.5 / n
    seg = 0.48 / 21
    wp = np.array([[-0.1, 0.25, 0.01 + (i + 0.5) * seg] for i in range(21)])
    wt = np.tile([0.0, 0.0, 1.0], (21, 1))
    k0 = W.k0_of(250e6)
    res = {}
    for ground, tag in ((0, "free"), (1, "pec")):
        centers = W.cube_cells(n, pitch) + np.array([0.0, 0.0, 0.05 if ground else 0.0])
        sizes = np.full_like(centers, pitch)
        pairs = len(wp) * len(centers)
        ms = vie.time_volume_conductor_rows(k0, centers, sizes, wp, wt, ground=ground, repeats=args.steps)
        t = time.perf_counter()
        vie.volume_conductor_rows(k0, centers, sizes, wp, wt, ground=ground)
        t_call = time.perf_counter() - t
        res[tag] = {"kernel_ms": ms, "pairs": pairs, "pairs_per_s": pairs / (ms * 1e-3), "call_ms": t_call * 1e3,
                    "row_bytes": pairs * 3 * 16}
        log(f"rows {tag}: {pairs} pairs, kernel {ms:.3f} ms ({pairs / (ms * 1e-3):.3e} pairs/s), "
            f"call with {pairs * 48 / 1e6:.0f} MB to the host {t_call * 1e3:.1f} ms")
    cpu = None
    if not args.no_cpu:
        from oracle import conductor_rows as R
        cores = min(16, len(os.sched_getaffinity(0)))
        centers = W.cube_cells(n, pitch)
        sample = centers[:: max(1, len(centers) // args.rows_cpu_cells)][: args.rows_cpu_cells]
        sizes = np.full_like(sample, pitch)
        t = time.perf_counter()
        R.conductor_rows_c(k0, sample, sizes, wp, wt, threads=cores)
        t_cpu = time.perf_counter() - t
        rate = len(wp) * len(sample) / t_cpu
        cpu = {"value": rate, "unit": "pairs/s", "cores": cores, "kind": "port",
               "sample": f"21 segments x {len(sample)} cells (free space) through oracle/conductor_rows.c "
                         f"(OpenMP, {cores} threads) in {t_cpu:.2f} s"}
        log(f"rows CPU port: {rate:.3e} pairs/s on {cores} threads")
    free = res["free"]
    # FP64 VALU issue: the wire-row z loop of k_conductor_rows<kSeries4> is 108
    # VALU instructions per two nodes in the AMD assembly of this build (54 a
    # node, one a quarter-rate v_rsq_f64); peak = 256 CUs x 4 SIMDs x 16 lanes
    # x 2.4 GHz lane-instructions/s (78.6 TFLOP/s FP64 counting an FMA as 2).
    per_node, peak = 54, 256 * 4 * 16 * 2.4e9
    achieved = free["pairs_per_s"] * 1000 * per_node
    roof = {"bound": "valu-fp64", "achieved": achieved, "peak": peak, "unit": "FP64 VALU lane-instructions/s",
            "frac": achieved / peak, "traffic": None, "kernel": "k_conductor_rows<kSeries4> (wire rows)",
            "instructions_per_node": per_node}
    log(f"rows roofline: {achieved:.3e} of {peak:.3e} FP64 VALU lane-instructions/s ({achieved / peak:.1%})")
    print(json.dumps({
        "metric": "volume-to-wire row pairs/s (10^3-point cuboid integral per segment-cell pair)",
        "value": free["pairs_per_s"], "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": free["kernel_ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{n}^3 cube (edge 0.5 m), 21-segment dipole 0.1 m from a face, 250 MHz; "
                               "the B rows of the coupled system", "cells": n ** 3, "segments": 21},
        "rows": res, "cpu_baseline": cpu, "roofline": roof,
    }))
    return 0


def bench_mesh(args):
    """The mesher's relationship checks on the per-voxel form of the sweep's
    model: one BOX record per voxel of the n^3 cube (edge 0.5 m) and the
    21-segment dipole (radius 1 mm) 0.1 m from one face.  The box-interior
    overlap scan of validate_box_relationships (dielectric_mesh.c:81-123)
    tests all n^3 (n^3 - 1) / 2 pairs; the clearance of evaluate_clearance
    (dielectric_mesh.c:215-316) every (box, segment) pair.  Kernel times with
    the boxes resident (HIP events), plus the whole calls.  CPU baseline: the
    C restatement oracle/mesh_check.c (the reference's pair loop without its
    per-pair region lookups, so faster than the reference) on one core, on
    the pairs of a bounded sample of left boxes, and the full clearance."""
    from nec2c_amd import vie
    n = args.mesh_size
    pitch = 0.5 / n
    idx = np.indices((n, n, n)).reshape(3, -1).T.astype(float)
    lo = idx * pitch
    hi = lo + pitch
    cell = np.full(len(lo), pitch)
    seg_len = 0.48 / 21
    zc = 0.01 + (np.arange(21) + 0.5) * seg_len
    seg = [np.full(21, -0.1), np.full(21, 0.25), zc, np.full(21, seg_len), np.zeros(21), np.zeros(21), np.ones(21),
           np.full(21, 1e-3)]
    B = len(lo)
    pairs = B * (B - 1) // 2
    pair, ms_overlap = vie.box_overlap(lo, hi, repeats=args.steps)
    r, ms_clear = vie.conductor_clearance(lo, hi, cell, segments=seg, repeats=args.steps)
    t = time.perf_counter()
    vie.box_overlap(lo, hi)
    vie.conductor_clearance(lo, hi, cell, segments=seg)
    call_ms = (time.perf_counter() - t) * 1e3
    rate = pairs / (ms_overlap * 1e-3)
    log(f"mesh {n}^3 per-voxel: {B} boxes, {pairs:.3e} pairs in {ms_overlap:.2f} ms ({rate:.3e} pairs/s), "
        f"first overlap {pair}; clearance {B * 21} pairs {ms_clear:.3f} ms, minimum gap {r['minimum_gap']:.4g} m; "
        f"both calls from the host {call_ms:.1f} ms")
    cpu = None
    if not args.no_cpu:
        from oracle import mesh_check as M
        K = min(args.mesh_cpu_lefts, B)
        t = time.perf_counter()
        M.box_overlap(lo, hi, 0, K)
        t_ov = time.perf_counter() - t
        sampled = K * B - K * (K + 1) // 2
        t = time.perf_counter()
        M.clearance(lo, hi, cell, segments=seg)
        t_cl = time.perf_counter() - t
        cpu_rate = sampled / t_ov
        cpu = {"value": cpu_rate, "unit": "pairs/s", "cores": 1, "kind": "port",
               "sample": f"the pairs of the first {K} left boxes ({sampled:.3e} pairs) through oracle/mesh_check.c "
                         f"in {t_ov:.2f} s; the full scan would take {pairs / cpu_rate:.0f} s; the full clearance "
                         f"({B * 21} pairs) {t_cl:.2f} s",
               "clearance_s": t_cl}
        log(f"mesh CPU port: {cpu_rate:.3e} pairs/s on one core (full scan {pairs / cpu_rate:.0f} s)")
    # LDS issue: every pair of k_box_overlap reads the right-hand box's six
    # coordinates from LDS as 3 broadcast ds_read2st64_b64, beside 20
    # VALU instructions (9 v_max_f64, 3 v_min_f64, 3 v_cmp_nlt_f64, a move and
    # the loop's 4 integer ops).  A build with 14 VALU instructions per pair ran
    # no faster, so the LDS reads bound the scan: one LDS instruction per CU per
    # cycle, peak = 256 CUs x 2.4 GHz wave-level LDS instructions/s, each pair
    # costing 3/64 of a wave-level read.  The VALU issue fraction is reported
    # beside it.
    lds_per_pair, peak = 3 / 64, 256 * 2.4e9
    achieved = rate * lds_per_pair
    valu_per_pair, valu_peak = 20, 256 * 4 * 16 * 2.4e9
    roof = {"bound": "lds", "achieved": achieved, "peak": peak, "unit": "wave-level LDS read instructions/s",
            "frac": achieved / peak, "traffic": None, "kernel": "k_box_overlap (all box pairs)",
            "lds_reads_per_pair": 3, "valu_instructions_per_pair": valu_per_pair,
            "valu_issue_frac": rate * valu_per_pair / valu_peak}
    log(f"mesh roofline: {achieved:.3e} of {peak:.3e} LDS read instructions/s ({achieved / peak:.1%}); "
        f"VALU issue {rate * valu_per_pair / valu_peak:.1%}")
    print(json.dumps({
        "metric": "box pairs/s of the mesher's interior-overlap scan (validate_box_relationships), per-voxel model",
        "value": rate, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": ms_overlap + ms_clear, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{n}^3 per-voxel BOX records (edge 0.5 m) + 21-segment dipole 0.1 m from a face",
                   "boxes": B, "pairs": pairs, "segments": 21},
        "mesh": {"overlap_ms": ms_overlap, "clearance_ms": ms_clear, "calls_ms": call_ms, "first_overlap": pair,
                 "minimum_gap_m": r["minimum_gap"], "required_gap_m": r["required_gap"],
                 "clearance_insufficient": r["insufficient"]},
        "cpu_baseline": cpu, "roofline": roof,
    }))
    return 0


def bench_columns(args):
    """Volume rows of the conductor columns (wire -> volume) of the sweep's
    coupled system: the n^3 cube (edge 0.5 m) and the 21-segment dipole 0.1 m
    from one face at 250 MHz, one column per unknown with NEC's straight-wire
    sparsity (segments j-1, j, j+1), NEC's near field at the 6^3 points of
    every cell, free space and over a PEC plane 0.05 m below the cube.
    Kernel time with the inputs resident (HIP events), plus the whole call
    with the columns copied to the host.  CPU baseline: the C restatement
    oracle/wire_volume.c with OpenMP over the (column, cell) tasks, as the
    reference runs them, on a bounded sample of cells."""
    from nec2c_amd import vie, workloads as W
    n = args.columns_size
    pitch = 0.5 / n
    f_mhz = 250.0
    wlam, k0 = 299.8 / f_mhz, W.k0_of(f_mhz * 1e6)
    rng = np.random.default_rng(5)
    src = np.zeros((21, 63), dtype=np.complex128)
    for j in range(21):
        for s in (j - 1, j, j + 1):
            if 0 <= s < 21:
                src[j, 3 * s:3 * s + 3] = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    res = {}
    for ground, tag in ((0, "free"), (1, "pec")):
        lift = 0.05 if ground else 0.0
        centers = W.cube_cells(n, pitch) + np.array([0.0, 0.0, lift])
        sizes = np.full_like(centers, pitch)
        chi = np.full(len(centers), 3.0 - 0.1j)
        seg = W.nec_straight_wire([-0.1, 0.25, 0.01 + lift], [-0.1, 0.25, 0.49 + lift], 21, 0.001, wlam)
        pairs = 21 * len(centers)
        ms = vie.time_conductor_volume_columns(k0, wlam, centers, sizes, chi, seg, src, ground=ground,
                                               repeats=args.steps)
        t = time.perf_counter()
        vie.conductor_volume_columns(k0, wlam, centers, sizes, chi, seg, src, ground=ground)
        t_call = time.perf_counter() - t
        res[tag] = {"kernel_ms": ms, "pairs": pairs, "pairs_per_s": pairs / (ms * 1e-3), "call_ms": t_call * 1e3}
        log(f"columns {tag}: {pairs} (column, cell) pairs, kernels {ms:.2f} ms ({pairs / (ms * 1e-3):.3e} pairs/s), "
            f"call with {pairs * 48 / 1e6:.0f} MB to the host {t_call * 1e3:.1f} ms")
    cpu = None
    if not args.no_cpu:
        cores = min(16, len(os.sched_getaffinity(0)))
        centers = W.cube_cells(n, pitch)
        sample = centers[:: max(1, len(centers) // args.columns_cpu_cells)][: args.columns_cpu_cells]
        from oracle import wire_volume as V
        dip = V.Conductors(*W.nec_straight_wire([-0.1, 0.25, 0.01], [-0.1, 0.25, 0.49], 21, 0.001, wlam))
        t = time.perf_counter()
        V.conductor_volume(k0, wlam, sample, np.full_like(sample, pitch), 3.0 - 0.1j, dip, src, threads=cores)
        t_cpu = time.perf_counter() - t
        rate = 21 * len(sample) / t_cpu
        cpu = {"value": rate, "unit": "pairs/s", "cores": cores, "kind": "port",
               "sample": f"21 columns x {len(sample)} cells (free space) through oracle/wire_volume.c "
                         f"(OpenMP, {cores} threads) in {t_cpu:.2f} s"}
        log(f"columns CPU port: {rate:.3e} pairs/s on {cores} threads")
    free = res["free"]
    print(json.dumps({
        "metric": "conductor-to-volume column pairs/s (NEC near field averaged over 6^3 points per column-cell pair)",
        "value": free["pairs_per_s"], "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": free["kernel_ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{n}^3 cube (edge 0.5 m), 21-segment dipole 0.1 m from a face, 250 MHz; "
                               "the D columns of the coupled system", "cells": n ** 3, "columns": 21},
        "columns": res, "cpu_baseline": cpu, "roofline": None,
    }))
    return 0


def bench_cross_slab(args):
    """The cross blocks of the 128^3 slab (1 cm voxels) and its 4-element Yagi
    (84 segments, the slab workload's geometry) at 200 MHz in free space: the
    84 conductor rows (10^3-point rule per pair) and the 84 conductor columns
    (NEC's near field at 216 points per cell, each element's unknowns with a
    straight wire's basis sparsity), kernel time with the inputs resident.
    CPU baseline: the C restatements on a bounded sample of cells, scaled per
    pair."""
    from nec2c_amd import vie, workloads as W
    n, h, f = args.slab_size, 0.01, 200e6
    k0, wlam = W.k0_of(f), 299.8 / (f / 1e6)
    edge = n * h
    yc, zc = 0.5 * edge, 0.5 * edge
    lengths = (0.50, 0.47, 0.44, 0.43)
    centers = W.cube_cells(n, h)
    sizes = np.full_like(centers, h)
    wires = [W.nec_straight_wire([-0.1, yc + (i - 1.5) * 0.2, zc - L / 2], [-0.1, yc + (i - 1.5) * 0.2, zc + L / 2],
                                 21, 1e-3, wlam) for i, L in enumerate(lengths)]
    seg = tuple(np.concatenate([w[q] for w in wires]) for q in range(8))
    nseg = len(seg[0])
    points = np.stack([seg[0], seg[1], seg[2]], -1) * wlam
    tangents = np.stack([seg[5], seg[6], seg[7]], -1)
    src = np.zeros((nseg, 3 * nseg), dtype=np.complex128)
    for j in range(nseg):
        for s_ in (j - 1, j, j + 1):
            if 0 <= s_ < nseg and s_ // 21 == j // 21:
                src[j, 3 * s_:3 * s_ + 3] = [1.0, 0.25j, -0.5]
    pairs = nseg * len(centers)
    log(f"cross blocks: {n}^3 cells x {nseg} segments = {pairs} pairs")
    ms_rows = vie.time_volume_conductor_rows(k0, centers, sizes, points, tangents, repeats=args.steps)
    log(f"rows: {ms_rows:.1f} ms ({pairs / (ms_rows * 1e-3):.3e} pairs/s)")
    chi = np.full(len(centers), 5.0 - 0.9j)
    ms_cols = vie.time_conductor_volume_columns(k0, wlam, centers, sizes, chi, seg, src, repeats=args.steps)
    log(f"columns: {ms_cols:.1f} ms ({pairs / (ms_cols * 1e-3):.3e} pairs/s)")
    cpu = None
    if not args.no_cpu:
        from oracle import conductor_rows as R, wire_volume as V
        cores = min(16, len(os.sched_getaffinity(0)))
        sample = centers[:: len(centers) // 512][:512]
        ss = np.full_like(sample, h)
        t = time.perf_counter()
        R.conductor_rows_c(k0, sample, ss, points, tangents, threads=cores)
        t_rows = time.perf_counter() - t
        t = time.perf_counter()
        V.conductor_volume(k0, wlam, sample, ss, 5.0 - 0.9j, V.Conductors(*seg), src, threads=cores)
        t_cols = time.perf_counter() - t
        per_cell = (t_rows + t_cols) / len(sample)
        cpu = {"value": 1.0 / (per_cell * len(centers)), "unit": "cross-block assemblies/s", "cores": cores,
               "kind": "port", "sample": f"{len(sample)} cells x {nseg} conductors through oracle/conductor_rows.c "
                                         f"({t_rows:.2f} s) and oracle/wire_volume.c ({t_cols:.2f} s), OpenMP "
                                         f"{cores} threads, scaled to {len(centers)} cells"}
        log(f"CPU port: rows {t_rows / len(sample) * len(centers):.0f} s + columns "
            f"{t_cols / len(sample) * len(centers):.0f} s for the full slab on {cores} threads")
    total = ms_rows + ms_cols
    print(json.dumps({
        "metric": "cross-block assemblies/s (conductor rows + columns of the coupled system)",
        "value": 1e3 / total, "unit": "assemblies/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": total, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{n}^3 slab (1 cm voxels), 4-element Yagi (84 segments) 0.1 m from a face, 200 MHz, "
                               "free space", "cells": n ** 3, "conductors": nseg},
        "rows_ms": ms_rows, "columns_ms": ms_cols, "pairs": pairs, "roofline": None, "cpu_baseline": cpu,
    }))
    return 0


def bench_point(args):
    """One frequency point of the 64^3 coupled sweep from scratch, all on the
    device except the conductor block C: prepare (cell map, Galerkin spectrum,
    FFTs), permittivity, C (the synthetic dipole matrix, host), the conductor
    rows B (dielectric_fft_assemble_conductor_rows) and columns D
    (dielectric_fft_assemble_conductor_columns, basis coefficients with a
    straight wire's sparsity) straight into the system, then the mixed solve
    to 1e-8.  `--steps` points of the 101-point sweep, each timed stage by
    stage with the device synchronised around it."""
    from nec2c_amd import vie, workloads as W
    n = args.size
    pitch = 0.5 / n
    eps = W.random_permittivity((n, n, n), seed=2).reshape(-1)
    centers = W.cube_cells(n, pitch)
    sizes = np.full_like(centers, pitch)
    dip = W.Dipole(center=(-0.1, 0.25, 0.25))
    seg_count = dip.segments
    points = dip.points()
    tangents = np.tile([0.0, 0.0, 1.0], (seg_count, 1))
    src = np.zeros((seg_count, 3 * seg_count), dtype=np.complex128)
    for j in range(seg_count):
        for s_, v in ((j - 1, 0.5), (j, 1.0), (j + 1, 0.5)):
            if 0 <= s_ < seg_count:
                src[j, 3 * s_:3 * s_ + 3] = [v, 0.25j * v, -0.5 * v]
    opts = dict(restart=args.restart, maximum_iterations=args.max_iterations, relative_tolerance=1e-8)
    stages = {k: [] for k in ("prepare", "conductor_c", "rows_b", "columns_d", "solve", "total")}
    iters = []
    for f in sweep_frequencies(args.steps, offset=0)[: args.steps]:
        k0 = W.k0_of(f)
        wlam = 299.8 / (f / 1e6)
        seg = W.nec_straight_wire(np.array(dip.center) - [0, 0, dip.length / 2],
                                  np.array(dip.center) + [0, 0, dip.length / 2], seg_count, dip.radius, wlam)
        vie.synchronize()
        t0 = time.perf_counter()
        s = vie.System(centers, [vie.MaterialSpec(1, 1.5)], seg_count, k0, pitch=pitch)
        s.set_permittivities(eps)
        vie.synchronize()
        t1 = time.perf_counter()
        s.set_blocks(C=dip.impedance_matrix(k0))
        t2 = time.perf_counter()
        s.assemble_rows(k0, centers, sizes, points, tangents)
        t3 = time.perf_counter()
        s.assemble_columns(k0, wlam, centers, sizes, eps - 1.0, seg, src)
        t4 = time.perf_counter()
        b = np.zeros(s.order, dtype=np.complex128)
        b[:seg_count] = dip.rhs()
        ok, x, rep = s.solve(b, precision=vie.MIXED, **opts)
        vie.synchronize()
        t5 = time.perf_counter()
        for k, v in zip(stages, (t1 - t0, t2 - t1, t3 - t2, t4 - t3, t5 - t4, t5 - t0)):
            stages[k].append(v)
        iters.append(rep.iterations)
        log(f"point {f / 1e6:.0f} MHz: prepare {1e3 * (t1 - t0):.0f} ms, C {1e3 * (t2 - t1):.0f} ms, "
            f"B rows {1e3 * (t3 - t2):.0f} ms, D columns {1e3 * (t4 - t3):.0f} ms, mixed solve {1e3 * (t5 - t4):.0f} ms "
            f"({rep.iterations} iterations, {'converged' if ok else 'NOT converged'}), total {t5 - t0:.3f} s")
        del s
    mean = {k: 1e3 * float(np.mean(v)) for k, v in stages.items()}
    print(json.dumps({
        "metric": "coupled frequency points/s from scratch (prepare + cross blocks + mixed solve to 1e-8)",
        "value": 1e3 / mean["total"], "unit": "points/s", "n_gpus": 1, "steps": args.steps, "warmup": 0,
        "ms_per_step": mean["total"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (assembly, outer solve) / f32 (inner cycles)", "data": "synthetic",
        "config": {"workload": f"{n}^3 random-permittivity cube (edge 0.5 m), 21-segment dipole 0.1 m from a face; "
                               "points of the 100-400 MHz sweep; C synthetic, B and D by the reference's formulas",
                   "cells": n ** 3, "conductors": seg_count},
        "stage_ms": mean, "iterations": iters, "roofline": None, "cpu_baseline": None,
    }))
    return 0


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed solves (slab: mixed solves of the slab, default 5; sweep: frequency points, "
                         "default 8)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed warm-up solves (default 1)")
    ap.add_argument("--fp64-steps", type=int, default=2,
                    help="slab: timed FP64 solves of each FP64 variant (after one warm-up)")
    ap.add_argument("--cpu-iterations", type=int, default=2,
                    help="slab: reference GMRES iterations timed for the CPU baseline")
    ap.add_argument("--free-space", action="store_true",
                    help="slab: the slab in free space instead of over the real ground")
    ap.add_argument("--geometry", choices=("sweep", "slab"), default="sweep",
                    help="matvec: the geometry whose fused apply is timed (counter runs)")
    ap.add_argument("--size", type=int, default=64, help="voxels per cube edge (64 = the 64^3 sweep)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU reference baseline")
    ap.add_argument("--matvec-repeats", type=int, default=50)
    ap.add_argument("--restart", type=int, default=120,
                    help="GMRES restart (30, the reference's fixed value, stagnates above ~250 MHz here)")
    ap.add_argument("--max-iterations", type=int, default=2000)
    ap.add_argument("--inner-reduction", type=float, default=0.0,
                    help="mixed: FP32 cycle stops at |g| <= max(target, this * beta); 0 = library default")
    ap.add_argument("--recycle", type=int, default=0,
                    help="mixed: recycled-subspace size of the GMRES cycles (0 = library default, -1 = off)")
    ap.add_argument("--workload", choices=("sweep", "batch", "serve", "matvec", "slab", "farfield", "rows", "columns",
                                           "point", "cross_slab", "mesh"),
                    default="slab",
                    help="slab: the 128^3 slab + Yagi (default, the headline); sweep: the 64^3 frequency "
                         "sweep; batch: 256 uneven-grid solves with "
                         "1-8 ports each; serve: Poisson-arrival request stream; matvec: fused operator "
                         "applies of one sweep point plus the copy calibration kernel (for counter runs)")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30, help="matvec: bytes of the calibration copy")
    ap.add_argument("--slab-size", type=int, default=128, help="slab: voxels per edge (128 = the 128^3 slab)")
    ap.add_argument("--farfield-size", type=int, default=64, help="farfield: voxels per cube edge")
    ap.add_argument("--rows-size", type=int, default=64, help="rows: voxels per cube edge")
    ap.add_argument("--rows-cpu-cells", type=int, default=16384, help="rows: cells in the CPU-baseline sample")
    ap.add_argument("--columns-size", type=int, default=64, help="columns: voxels per cube edge")
    ap.add_argument("--columns-cpu-cells", type=int, default=8192, help="columns: cells in the CPU-baseline sample")
    ap.add_argument("--mesh-size", type=int, default=64, help="mesh: voxels per cube edge (one BOX per voxel)")
    ap.add_argument("--mesh-cpu-lefts", type=int, default=2048,
                    help="mesh: left boxes whose pairs the CPU baseline scans")
    ap.add_argument("--pass-sweep", action="store_true",
                    help="matvec: also time every launch variant of the y and x passes (counter runs then cover them all)")
    ap.add_argument("--arnoldi-column", type=int, default=0,
                    help="matvec: also run the Arnoldi step kernels at this basis column (0: skip)")
    ap.add_argument("--batch-count", type=int, default=256, help="batch: number of models")
    ap.add_argument("--inflight", type=int, default=96, help="batch: jobs on the device at once")
    ap.add_argument("--batch-cpu-jobs", type=int, default=24, help="batch: jobs in the CPU-baseline sample")
    ap.add_argument("--batch-kh", type=float, default=0.1,
                    help="batch: k0 sqrt(eps_r) h of every model (the workload bounds it by 0.5)")
    ap.add_argument("--requests", type=int, default=100, help="serve: number of requests")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: mean arrival rate (requests/s)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 5 if args.workload == "slab" else 8
    if args.warmup is None:
        args.warmup = 1
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        log("only --gpus 1 is supported: the benchmark is a single-GPU workload")
        return 2

    from nec2c_amd import vie, workloads as W
    if vie.device_count() < 1:
        log("no HIP device visible")
        return 1
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "matvec":
        return bench_matvec(args)
    if args.workload == "slab":
        return bench_slab(args)
    if args.workload == "sweep":
        return bench_sweep(args)
    if args.workload == "farfield":
        return bench_farfield(args)
    if args.workload == "rows":
        return bench_rows(args)
    if args.workload == "mesh":
        return bench_mesh(args)
    if args.workload == "columns":
        return bench_columns(args)
    if args.workload == "point":
        return bench_point(args)
    if args.workload == "cross_slab":
        return bench_cross_slab(args)
    return 2


def bench_sweep(args):
    """The 64^3 random-permittivity sweep (module docstring)."""
    from nec2c_amd import vie, workloads as W
    n = args.size
    pitch = 0.5 / n
    eps = W.random_permittivity((n, n, n), seed=2)
    centers = W.cube_cells(n, pitch)
    dipole = W.Dipole(center=(-0.1, 0.25, 0.25))
    freqs = sweep_frequencies(args.steps + args.warmup)
    opts = dict(restart=args.restart, maximum_iterations=args.max_iterations, relative_tolerance=1e-8)
    log(f"setting up {len(freqs)} frequency points at {n}^3 (c = {dipole.segments})")
    systems, cases = [], []
    t0 = time.perf_counter()
    for f in freqs:
        s, case = build_case(n, f, eps, dipole, None, centers, pitch, opts)
        systems.append(s)
        cases.append(case)
        log(f"  {f / 1e6:.1f} MHz ready ({time.perf_counter() - t0:.1f} s)")
    warm, timed = systems[: args.warmup], systems[args.warmup:]

    for s in warm:
        s.solve_resident(vie.MIXED, args.inner_reduction, args.recycle, **opts)
        s.solve_resident(vie.FP64, **opts)
    vie.synchronize()

    def run(prec, recycle=0):
        lat, reps, sols = [], [], []
        vie.synchronize()
        t = time.perf_counter()
        for s in timed:
            a = time.perf_counter()
            ok, rep = s.solve_resident(prec, args.inner_reduction, recycle, **opts)
            lat.append(time.perf_counter() - a)
            if not ok:
                raise RuntimeError(f"solve failed: {vie.STATUS[rep.status]} rel {rep.relative_residual}")
            reps.append(rep)
        vie.synchronize()
        return time.perf_counter() - t, lat, reps

    log("timing mixed-precision solves")
    for q in timed:
        q.solve_stats(reset=True)
    t_mixed, lat_mixed, rep_mixed = run(vie.MIXED, args.recycle)
    rec_k = timed[0].recycled
    col_stats = [q.solve_stats(reset=True) for q in timed]
    mean_col = sum(m * c for m, c in col_stats) / max(sum(c for _, c in col_stats), 1)
    x_mixed = [s.download_solution() for s in timed]
    log("timing FP64 solves")
    t_fp64, lat_fp64, rep_fp64 = run(vie.FP64)
    x_fp64 = [s.download_solution() for s in timed]
    # the same recycled cycles in FP64, to separate the precision gain from the recycling gain
    k_fp64r = args.recycle if args.recycle > 0 else 16
    log(f"timing FP64 solves with a recycled subspace of {k_fp64r}")
    for q in timed:  # untimed: sizes each basis for FP64 plus the recycled columns
        q.solve_resident(vie.FP64, 0.0, k_fp64r, **dict(opts, relative_tolerance=1.0))
    t_fp64r, _, rep_fp64r = run(vie.FP64, k_fp64r)

    # parity between the two precisions (feed impedance, conductor currents)
    fi = dipole.feed_index()
    z_err = max(abs(1 / xm[fi] - 1 / xd[fi]) / abs(1 / xd[fi]) for xm, xd in zip(x_mixed, x_fp64))
    i_err = max(float(np.max(np.abs(xm[:21] - xd[:21])) / np.max(np.abs(xd[:21])))
                for xm, xd in zip(x_mixed, x_fp64))
    gain_db = max(gain_difference_db(xm, xd, case[4], centers, pitch, dipole.points(), dipole.delta)
                  for xm, xd, case in zip(x_mixed, x_fp64, cases[args.warmup:]))
    log(f"far-field gain, mixed vs FP64: max {gain_db:.2e} dB over a 5-degree grid")

    # matvec timing and roofline of the dominant kernel
    s0 = timed[0]
    variants = s0.tune()
    mv_ms32, pass32 = s0.time_apply(vie.MIXED, args.matvec_repeats)
    mv_ms64, pass64 = s0.time_apply(vie.FP64, args.matvec_repeats)
    bytes32, pbytes32 = s0.apply_bytes(vie.MIXED)
    bytes64, pbytes64 = s0.apply_bytes(vie.FP64)
    names = ["zfwd", "yfwd", "xpass", "yinv", "zinv", "coupling"]
    dom = int(np.argmax(pass32[:5]))
    achieved = pbytes32[dom] / (pass32[dom] * 1e-3) / 1e9

    # the two streaming kernels of an Arnoldi step (the solve's dominant
    # kernels), timed at the mean basis column of the timed mixed solves
    arn = {}
    jbar = max(1, int(round(mean_col)))
    for prec, tag in ((vie.MIXED, "fp32"), (vie.FP64, "fp64")):
        t = s0.time_arnoldi(prec, jbar, args.matvec_repeats)
        t["gbs_reduce"] = t["bytes_reduce"] / (t["ms_reduce"] * 1e-3) / 1e9
        t["gbs_update"] = t["bytes_update"] / (t["ms_update"] * 1e-3) / 1e9
        arn[tag] = t

    iters_mixed = [r.iterations for r in rep_mixed]
    apps_mixed = sum(r.operator_applications for r in rep_mixed)
    apps_fp64 = sum(r.operator_applications for r in rep_fp64)
    K = len(timed)
    result = {
        "metric": METRIC,
        "value": K / t_mixed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_mixed / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 inner (fp32 complex) / c128 outer (fp64 complex)",
        "data": "synthetic",
        "config": {
            "workload": f"{n}^3 random-eps lossy cube (edge 0.5 m) + 21-segment dipole at 0.1 m, "
                        f"{K} points of the 100-400 MHz sweep, GMRES({args.restart}) to 1e-8",
            "grid": [n, n, n], "padded": [2 * n, 2 * n, 2 * n], "conductors": dipole.segments,
            "unknowns": s0.order, "frequencies_mhz": [round(f / 1e6, 2) for f in freqs[args.warmup:]],
        },
        "fp64_solves_per_s": K / t_fp64,
        "speedup_mixed_vs_fp64": t_fp64 / t_mixed,
        "recycled_subspace_mixed": rec_k,
        "fp64_recycled_solves_per_s": K / t_fp64r,
        "speedup_mixed_vs_fp64_recycled": t_fp64r / t_mixed,
        "iterations_fp64_recycled": [r.iterations for r in rep_fp64r],
        "matvecs_per_s_mixed": 1e3 / mv_ms32,
        "matvecs_per_s_fp64": 1e3 / mv_ms64,
        "solve_matvecs_per_s_mixed": apps_mixed / t_mixed,
        "solve_matvecs_per_s_fp64": apps_fp64 / t_fp64,
        "p50_ms_mixed": 1e3 * float(np.percentile(lat_mixed, 50)),
        "p99_ms_mixed": 1e3 * float(np.percentile(lat_mixed, 99)),
        "p50_ms_fp64": 1e3 * float(np.percentile(lat_fp64, 50)),
        "p99_ms_fp64": 1e3 * float(np.percentile(lat_fp64, 99)),
        "iterations_mixed": iters_mixed,
        "iterations_fp64": [r.iterations for r in rep_fp64],
        "residual_max_mixed": max(r.relative_residual for r in rep_mixed),
        "residual_max_fp64": max(r.relative_residual for r in rep_fp64),
        "feed_impedance_rel_err_mixed_vs_fp64": z_err,
        "conductor_current_rel_err_mixed_vs_fp64": i_err,
        "far_field_gain_max_db_mixed_vs_fp64": gain_db,
        "feed_impedance_ohm_fp64": [str(complex(np.round(1 / x[fi], 6))) for x in x_fp64[:3]],
        "matvec": {
            "ms_mixed": mv_ms32, "ms_fp64": mv_ms64,
            "bytes_mixed": bytes32, "bytes_fp64": bytes64,
            "effective_gbs_mixed": bytes32 / (mv_ms32 * 1e-3) / 1e9,
            "effective_gbs_fp64": bytes64 / (mv_ms64 * 1e-3) / 1e9,
            "pass_ms_mixed": dict(zip(names, pass32)),
            "pass_ms_fp64": dict(zip(names, pass64)),
            "pass_bytes_mixed": dict(zip(names, pbytes32)),
            "tuned_variants": variants,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_dots_dc<float> (delayed-CGS2 reduction; basis column {jbar}, the mean over the "
                      f"timed mixed solves; ({jbar} + 2) x {s0.order} complex float read per launch)",
            "achieved": arn["fp32"]["gbs_reduce"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": arn["fp32"]["gbs_reduce"] / HBM_PEAK_GBS,
            "traffic": None,
        },
        "roofline_matvec": {
            "bound": "hbm",
            "kernel": pass_kernel_name(dom, 2 * n, variants["fp32"][dom]),
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
        },
        "cpu_baseline": None,
        "arnoldi": {"j": jbar, "mean_column": mean_col, **arn},
    }
    traffic = profiled_traffic(result["roofline_matvec"]["kernel"])
    if traffic is not None:
        result["roofline_matvec"]["traffic"] = traffic
    traffic = profiled_traffic(f"k_dots_dc<float>@{jbar}")
    if traffic is not None:
        result["roofline"]["traffic"] = traffic
    if not args.no_cpu:
        log("CPU baseline: reference FFT-GMRES on one frequency point")
        cb, check = cpu_baseline(n, pitch, eps, cases[args.warmup], rep_fp64[0].iterations, args.restart,
                                 x_sols={"mixed": x_mixed[0], "fp64": x_fp64[0]}, feed=fi,
                                 gain_args=(centers, pitch, dipole.points(), dipole.delta))
        result["cpu_baseline"] = cb
        # the FP64 oracle (the reference's own operator and GMRES) on the first timed point
        result["oracle_check"] = check
    print(json.dumps(result), flush=True)
    for s in systems:
        s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
