#!/usr/bin/env python3
"""Benchmark of the MI355X VIE FFT-GMRES on the 64^3 random-permittivity sweep.

Workload (one "step" = one frequency point solved to a 1e-8 true residual):
64 x 64 x 64 voxel block, edge 0.5 m, eps_r per voxel from a seeded Gaussian
random field (correlation 4 voxels) mapped to [1.5, 10], loss tangent U[0, 0.05];
a z-directed 0.48 m, 21-segment, 1 mm dipole with a 1 V centre delta gap, 0.1 m
from the x = 0 face; frequencies from the 101-point 100-400 MHz sweep.  The
operator is the reference's coupled [C B; D V] (c = 21, v = 786432); C/B/D are
synthetic blocks of the reference's shape (nec2c_amd/workloads.py).

Reported: mixed-precision (FP32 cycles, FP64 true residual) solves/s as the
headline value, FP64 solves/s beside it, matvecs/s, p50/p99 solve latency, the
roofline of the dominant kernel (X pass) from HIP events on the system stream,
and the reference's own CPU FFT-GMRES (oracle/_ref) on a bounded sample.

All inputs (spectrum, permittivity, C/B/D, right-hand side) are resident in
HBM before the timed region; each timed solve starts from x0 = 0.
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


def pass_kernel_name(pas, P, v, E=4096):
    """Kernel instantiation of FP32 volume pass `pas` at launch variant v
    (lines per workgroup and workgroup size as the library's tile rules)."""
    cap = (8, 16, 16, 16, 8)[pas]
    l0 = min(E // P, cap)
    lw = (l0, max(l0 // 2, 1), max(l0 // 4, 1), max(l0 // 2, 1))[v]
    wg = (256, 128, 64, 256)[v]
    kern = ("k_zfwd<float,{P},{lw},{wg}>", "k_ypass<float,{P},{lw},{wg},false>",
            "k_xpass<float,{P},{lw},{wg}> (X FFT + 3x3 spectral multiply + inverse X)",
            "k_ypass<float,{P},{lw},{wg},true>", "k_zinv<float,{P},{lw},{wg}>")[pas]
    return kern.format(P=P, lw=lw, wg=wg)


def cpu_baseline(n, pitch, eps, case, iterations_needed, restart, sample_iters=10):
    """The reference's own code (oracle/_ref): vie_fft_create + vie_gmres on the
    same coupled system for one frequency point, bounded to ``sample_iters``
    GMRES iterations and scaled to a full solve by the iteration count the GPU
    FP64 solve (the same algorithm) needed at that point.  The scaling is
    optimistic for the CPU: later Arnoldi steps orthogonalise against more
    vectors than the sampled first ones."""
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
    total = t_build + per_it * iterations_needed
    return {
        "value": 1.0 / total,
        "unit": "solves/s",
        "cores": cores,
        "kind": "reference",
        "sample": (f"1 frequency point ({k0 * 299792458.0 / 2 / math.pi / 1e6:.1f} MHz) of the same {n}^3 coupled "
                   f"system: reference vie_fft_create (builtin FFT, OpenMP) {t_build:.2f} s + the first "
                   f"{rep.iterations} vie_gmres({restart}) iterations {t_iters:.2f} s, scaled to the "
                   f"{iterations_needed} iterations the same algorithm needs there; one matvec {t_mv * 1e3:.1f} ms; "
                   "[C B; D] products restated around the reference volume operator (its prepare caps at 512 MiB)"),
        "matvec_ms": t_mv * 1e3,
        "build_s": t_build,
        "seconds_per_iteration": per_it,
        "iterations_sampled": rep.iterations,
        "iterations_scaled_to": iterations_needed,
    }, None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed frequency points (solves) per precision")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=64, help="voxels per cube edge (64 = the 64^3 sweep)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU reference baseline")
    ap.add_argument("--matvec-repeats", type=int, default=50)
    ap.add_argument("--restart", type=int, default=120,
                    help="GMRES restart (30, the reference's fixed value, stagnates above ~250 MHz here)")
    ap.add_argument("--max-iterations", type=int, default=2000)
    ap.add_argument("--inner-reduction", type=float, default=0.0,
                    help="mixed: FP32 cycle stops at |g| <= max(target, this * beta); 0 = library default")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        log("only --gpus 1 is supported: the benchmark is a single-GPU workload")
        return 2

    from nec2c_amd import vie, workloads as W
    if vie.device_count() < 1:
        log("no HIP device visible")
        return 1

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
        s.solve_resident(vie.MIXED, args.inner_reduction, **opts)
        s.solve_resident(vie.FP64, **opts)
    vie.synchronize()

    def run(prec):
        lat, reps, sols = [], [], []
        vie.synchronize()
        t = time.perf_counter()
        for s in timed:
            a = time.perf_counter()
            ok, rep = s.solve_resident(prec, args.inner_reduction, **opts)
            lat.append(time.perf_counter() - a)
            if not ok:
                raise RuntimeError(f"solve failed: {vie.STATUS[rep.status]} rel {rep.relative_residual}")
            reps.append(rep)
        vie.synchronize()
        return time.perf_counter() - t, lat, reps

    log("timing mixed-precision solves")
    t_mixed, lat_mixed, rep_mixed = run(vie.MIXED)
    x_mixed = [s.download_solution() for s in timed]
    log("timing FP64 solves")
    t_fp64, lat_fp64, rep_fp64 = run(vie.FP64)
    x_fp64 = [s.download_solution() for s in timed]

    # parity between the two precisions (feed impedance, conductor currents)
    fi = dipole.feed_index()
    z_err = max(abs(1 / xm[fi] - 1 / xd[fi]) / abs(1 / xd[fi]) for xm, xd in zip(x_mixed, x_fp64))
    i_err = max(float(np.max(np.abs(xm[:21] - xd[:21])) / np.max(np.abs(xd[:21])))
                for xm, xd in zip(x_mixed, x_fp64))

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
            "kernel": pass_kernel_name(dom, 2 * n, variants["fp32"][dom]),
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        log("CPU baseline: reference FFT-GMRES on one frequency point")
        cb, _ = cpu_baseline(n, pitch, eps, cases[args.warmup], rep_fp64[0].iterations, args.restart)
        result["cpu_baseline"] = cb
    print(json.dumps(result), flush=True)
    for s in systems:
        s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
