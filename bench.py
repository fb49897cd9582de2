"""Benchmark: seamless-clone Poisson blend to 1e-6 relative residual on one GPU.

Workload (one step = one full solve through the C ABI, inputs resident in HBM):
4096x4096 RGBA f32, fbm source/target (frequency 0.125, 6 octaves, gain 0.5,
seeds 101+c / 201+c, mapped to [0, 1]), alpha 1, seeded 12-vertex star mask
covering ~30% of the image; R, G, B solved independently (alpha has a zero
initial residual and passes through at iteration 0, as in the reference).

Both forms run on the same inputs: the plain start (x = anchor) and the learned
warm start (paintop_amd/weights, raced against the plain start on the solver's
own residual). Prints ONE JSON line: the learned form's ms per megapixel to
tolerance (lower is better) with the plain form's beside it, the sweep kernel's
roofline, and the CPU baselines timed on this host.
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

METRIC = "ms/megapixel to 1e-6 rel. residual, learned warm-start vs plain SOR; HBM GB/s"
HBM_PEAK_GBPS = 8000.0
BYTES_PER_UNKNOWN_SWEEP = 12.0  # FP32 e read + e write + rhs read


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--size", type=int, default=4096)
    p.add_argument("--half-sweeps", type=int, default=0)
    p.add_argument("--omega", type=float, default=0.0)
    p.add_argument("--tile-rows", type=int, default=0)
    p.add_argument("--max-iterations", type=int, default=100000)
    p.add_argument("--tol", type=float, default=1e-6)
    p.add_argument("--cpu-seconds", type=float, default=20.0,
                   help="budget of the CPU baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--plain-only", action="store_true",
                   help="time only the plain start (no learned warm start)")
    p.add_argument("--learned-only", action="store_true",
                   help="time only the learned form (a profiling aid: prints its time and "
                        "sweep counts, with no plain comparison or roofline)")
    p.add_argument("--learned", action="store_true",
                   help="batch workload: run the learned warm start and corrections")
    p.add_argument("--workload", choices=["seamless", "screened", "batch", "outside-scan", "copy"],
                   default="seamless",
                   help="seamless: 4096^2 RGBA blend to 1e-6 (the headline); screened: "
                        "8192^2 Gray screened Poisson, lambda 0.01, to 1e-7; batch: 512 "
                        "uneven single-channel blends (sides 32..1024) in one call, to 1e-6; "
                        "copy: a 1 GiB device-to-device copy (the HBM reference point and the "
                        "rocprofv3 counter calibration)")
    return p.parse_args()


def cpu_baselines(size, omega, gpu_sweeps, budget_s):
    """Time the oracle on this host on a bounded sample of the same workload
    (the full-size R channel):
    * the OpenMP red-black SOR stand-in for cpu.optimized (the reference has
      no optimized Poisson kernel) on every CPU this process may run on, at the
      GPU's omega, timed over whole sweeps and scaled to the sweeps the plain
      GPU solve needs x 3 channels;
    * the single-core restatement of cpu.reference (lexicographic SOR, omega
      1.8, a full residual after every sweep) timed over 40 full-size sweeps,
      with its sweep count to 1e-6 projected (it cannot be run there: ~10^5..10^6
      sweeps, beyond the reference's own 100000 cap): Young's spectral radius of
      lexicographic SOR at omega 1.8 from the domain's smallest eigenvalue (the
      GPU tuner's, recovered from its omega), scaled by the ratio of actual to
      asymptotic sweeps measured on the same workload at 384^2 run to
      tolerance; and an upper estimate anchored at the measured residual after
      the 40 sweeps (tests/test_oracle_tight.py pins both projections)."""
    from oracle import oracle as orc
    from paintop_amd.inputs import seamless_clone_case
    from tests.helpers import classify, laplacian

    src, tgt, msk = seamless_clone_case(size, channels=4)
    s = src.numpy()[..., 0].copy()
    t = tgt.numpy()[..., 0].astype(np.float64)
    m = msk.numpy()
    for a in (src, tgt, msk):
        a.free()
    cells = classify(m)
    rhs = np.where(cells == 1, laplacian(s), 0.0)
    avail = len(os.sched_getaffinity(0))
    threads = avail
    # all-cores stand-in: a short run to size the sample, then the timed sample
    n = 4
    t0 = time.perf_counter()
    _, rep, used = orc.rbsor_omp(cells, rhs, t, 0.0, n, 1e-30, omega, threads=threads,
                                  stall_window=0)
    per_sweep = (time.perf_counter() - t0) / n
    n2 = max(2, min(400, int(budget_s / max(per_sweep, 1e-6))))
    t0 = time.perf_counter()
    _, rep, used = orc.rbsor_omp(cells, rhs, t, 0.0, n2, 1e-30, omega, threads=threads,
                                  stall_window=0)
    per_sweep = (time.perf_counter() - t0) / n2
    mp = size * size / 1e6
    opt_ms_per_mp = per_sweep * gpu_sweeps * 3 * 1e3 / mp
    # cpu.reference: 40 full-size sweeps on one core
    k0 = 40
    t0 = time.perf_counter()
    _, rref = orc.solve(cells, rhs, t, 0.0, k0, 1e-30, 1.8)
    ref_sweep = (time.perf_counter() - t0) / k0
    mu = orc.mu_from_optimal_omega(omega)
    rho = orc.sor_spectral_radius(mu, 0.0, 1.8)
    tol = 1e-6
    upper = k0 + math.log(tol / rref.residual_history[k0 - 1]) / math.log(rho)
    # the actual / asymptotic sweep ratio of the same workload at 384^2
    from tests.helpers import smallest_eigenvalue, star_case
    c3, r3, t3 = star_case(384)
    _, r384 = orc.solve(c3, r3, t3, 0.0, 100000, tol, 1.8)
    ratio = r384.iterations / orc.project_sor_sweeps(smallest_eigenvalue(c3), 0.0, 1.8, tol)
    central = orc.project_sor_sweeps(mu, 0.0, 1.8, tol) * ratio
    return {
        "value": round(opt_ms_per_mp, 3),
        "unit": "ms/MP",
        "cores": int(used),
        "kind": "port",
        "sample": (f"OpenMP red-black SOR (oracle/rbsor_omp.c, the cpu.optimized stand-in) on "
                   f"the full {size}x{size} R channel, {n2} sweeps timed on {used} threads "
                   f"({per_sweep * 1e3:.1f} ms/sweep incl. residual), scaled to the plain GPU "
                   f"solve's {gpu_sweeps} sweeps x 3 channels"),
        "host_cpus": {"os_cpu_count": os.cpu_count(), "affinity": avail,
                      "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")},
        "cpu_reference": {
            "kind": "port (oracle/poisson_oracle.c or_solve: lexicographic SOR, omega 1.8)",
            "cores": 1,
            "ms_per_sweep_full_image": round(ref_sweep * 1e3, 1),
            "sweeps_timed": k0,
            "rel_residual_after_timed": rref.residual_history[k0 - 1],
            "projected_sweeps": int(round(central)),
            "projected_sweeps_upper": int(round(upper)),
            "projected_ms_per_mp": round(central * ref_sweep * 3 * 1e3 / mp, 1),
            "projected_ms_per_mp_upper": round(upper * ref_sweep * 3 * 1e3 / mp, 1),
            "note": ("projected, not run: the reference stops at its 100000-sweep cap "
                     "(poisson_blend.rs:71) as max-iterations long before 1e-6; sweeps = "
                     "log(1e-6)/log(rho_SOR(1.8)) x the actual/asymptotic ratio measured at "
                     f"384^2 ({ratio:.3f}); upper = anchored at the residual after {k0} "
                     "measured sweeps"),
        },
    }


def bench_outside_scan(args, info):
    """assert.no_change_outside_mask@1's device scan after the blend: the
    4096^2 RGBA target against the solved candidate with the star mask (nothing
    leaks, so every pixel is read and no location is written). One step = one
    scan; HBM-bound: (2 x channels + 1) x 4 B per pixel."""
    from paintop_amd import hip
    from paintop_amd.device import DeviceArray, synchronize
    from paintop_amd.inputs import seamless_clone_case

    n, c = args.size, 4
    src, tgt, msk = seamless_clone_case(n, channels=c)
    out = DeviceArray(tgt.shape, np.float32)
    hip.poisson_solve_device(hip.make_desc(n, n, c, args.max_iterations, args.tol), src.ptr,
                             tgt.ptr, msk.ptr, out.ptr)

    def step():
        return hip.assert_no_change_outside_mask_device(tgt.ptr, out.ptr, msk.ptr, n, n, c)

    for _ in range(max(args.warmup, 1)):
        step()
    synchronize()
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        r = step()  # returns after the result is on the host
        times.append((time.perf_counter() - t0) * 1e3)
    ms = float(np.median(times))
    alg = (2 * c + 1) * 4.0 * n * n
    gbps = alg / (ms * 1e-3) / 1e9
    line = {
        "metric": METRIC, "value": round(ms / (n * n / 1e6), 5), "unit": "ms/MP", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (device-generated)",
        "config": {"workload": f"assert.no_change_outside_mask scan, {n}x{n} RGBA target vs "
                               f"solved candidate, star mask, encoded space",
                   "changed_outside": r["changed_outside"], "device": info["name"]},
        "roofline": {"kernel": "k_os_scan + k_os_total + k_os_locs (one scan, host round trip "
                               "included)", "bound": "hbm", "achieved": round(gbps, 1),
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(gbps / HBM_PEAK_GBPS, 4),
                     "traffic": None},
    }
    print(json.dumps(line), flush=True)


def bench_copy(args, info):
    """A 1 GiB device-to-device copy (hipMemcpy, the runtime's copy kernel):
    the achievable HBM rate next to the 8 TB/s peak, and a kernel of known
    traffic (1 GiB read + 1 GiB written per copy) for scaling rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters."""
    from paintop_amd.device import D2D, DeviceArray, runtime, synchronize

    nbytes = 1 << 30
    a = DeviceArray((nbytes // 4,), np.float32)
    b = DeviceArray((nbytes // 4,), np.float32)
    rt = runtime()

    def step():
        rc = rt.hipMemcpy(b.ptr, a.ptr, nbytes, D2D)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed: {rc}")

    for _ in range(max(args.warmup, 1)):
        step()
    synchronize()
    times = []
    for _ in range(args.steps):
        synchronize()
        t0 = time.perf_counter()
        step()
        synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
    ms = float(np.median(times))
    gbps = 2.0 * nbytes / (ms * 1e-3) / 1e9
    line = {
        "metric": "device-to-device copy GB/s (read + write)", "value": round(gbps, 1),
        "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "hipMemcpy device-to-device, 1 GiB", "device": info["name"]},
        "roofline": {"kernel": "__amd_rocclr_copyBuffer", "bound": "hbm",
                     "achieved": round(gbps, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(gbps / HBM_PEAK_GBPS, 4), "traffic": 2 * nbytes},
    }
    a.free()
    b.free()
    print(json.dumps(line), flush=True)


def bench_batch(args, info):
    from paintop_amd import hip
    from paintop_amd.device import DeviceArray, synchronize
    from paintop_amd.inputs import batch_case

    cases = batch_case(512, seed0=500)
    outs = [DeviceArray((h, w), np.float32) for (w, h, *_r) in cases]
    items = [(w, h, 1, s.ptr, t.ptr, m.ptr, o.ptr) for (w, h, s, t, m), o in zip(cases, outs)]
    flags = hip.PH_FLAG_TIMING
    if args.learned:
        from paintop_amd import warmstart
        warmstart.enable()
        flags |= hip.PH_FLAG_WARM_START_NET
    desc = hip.make_desc(0, 0, 1, args.max_iterations, args.tol, 0.0, hip.PH_OP_POISSON_BLEND,
                         omega=args.omega, half_sweeps=args.half_sweeps,
                         flags=flags, tile_rows=args.tile_rows)
    for _ in range(args.warmup):
        hip.poisson_solve_batch_device(desc, items, want_history=False)
    times, results = [], []
    for _ in range(args.steps):
        synchronize()
        t0 = time.perf_counter()
        res = hip.poisson_solve_batch_device(desc, items, want_history=False)
        synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
        results.append(res)
    ms = float(np.median(times))
    px = sum(w * h for (w, h, *_r) in cases)
    res = results[len(results) // 2]
    its = np.array([r.iterations for r in res])
    usw = sum(r.unknown_sweeps for r in res)
    sweep_ms = res[0].sweep_ms
    # compulsory traffic: 12 B per unknown per launch, each launch running K/2
    # sweeps per HBM round trip (as for the single-image workloads)
    K = max(res[0].half_sweeps, 2)
    achieved = BYTES_PER_UNKNOWN_SWEEP * usw / (K / 2) / (sweep_ms * 1e-3) / 1e9 \
        if sweep_ms > 0 else 0.0
    effective = BYTES_PER_UNKNOWN_SWEEP * usw / (sweep_ms * 1e-3) / 1e9 if sweep_ms > 0 else 0.0
    line = {
        "metric": METRIC, "value": round(ms / (px / 1e6), 4), "unit": "ms/MP", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 sweeps, f64 residual/refinement", "data": "synthetic (device-generated)",
        "config": {"workload": f"512 uneven single-channel fbm blends, sides log-uniform "
                               f"32..1024, ellipse/polygon masks 20-60%, tol {args.tol:g}, one "
                               f"batched call", "megapixels": round(px / 1e6, 2),
                   "sweeps_min": int(its.min()), "sweeps_median": int(np.median(its)),
                   "sweeps_max": int(its.max()), "launches": res[0].launches,
                   "all_converged": all(r.converged for r in res),
                   "form": "learned warm start" if args.learned else "plain",
                   "warm_start_kept": sum(r.warm_start_used for r in res),
                   "coarse_corrections": sum(r.coarse_corrections for r in res),
                   "device": info["name"]},
        "roofline": {"kernel": f"k_sor<{K},3> + k_sor_control<{K}>", "bound": "hbm",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
                     "effective_GBps_12B_per_unknown_sweep": round(effective, 1),
                     "note": "achieved = compulsory bytes (12 B per unknown per launch) / "
                             "hipEvent time of the sweep launches"},
    }
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the hip backend runs on one GPU; --gpus must be 1 "
                          "(sharding across GPUs is designed, not implemented)"}))
        sys.exit(2)
    from paintop_amd import hip
    from paintop_amd.device import DeviceArray, synchronize
    from paintop_amd.inputs import seamless_clone_case

    info = hip.probe()
    size = args.size
    if args.workload == "batch":
        return bench_batch(args, info)
    if args.workload == "outside-scan":
        return bench_outside_scan(args, info)
    if args.workload == "copy":
        return bench_copy(args, info)
    if args.workload == "screened":
        from paintop_amd.inputs import screened_case
        if size == 4096:
            size = 8192
        tol = 1e-7 if args.tol == 1e-6 else args.tol
        lam, chans, op = 0.01, 1, hip.PH_OP_SCREENED_POISSON
        src, tgt, msk = screened_case(size)
        wl = (f"screened Poisson {size}x{size} Gray f32, lambda 0.01 toward fbm, guidance fbm "
              f"+ N(0, 0.01), full image minus a 64 px border, tol {tol:g}")
    else:
        tol, lam, chans, op = args.tol, 0.0, 4, hip.PH_OP_POISSON_BLEND
        src, tgt, msk = seamless_clone_case(size, channels=4)
        wl = (f"seamless-clone Poisson blend {size}x{size} RGBA f32, star mask ~30%, fbm "
              f"source/target, tol {tol:g}")
    out = DeviceArray(tgt.shape, np.float32)
    stream = 0

    def timed(flags, steps=None, warmup=None):
        desc = hip.make_desc(size, size, chans, args.max_iterations, tol, lam, op,
                             omega=args.omega, half_sweeps=args.half_sweeps,
                             flags=hip.PH_FLAG_TIMING | flags, tile_rows=args.tile_rows)

        def step():
            return hip.poisson_solve_device(desc, src.ptr, tgt.ptr, msk.ptr, out.ptr, stream,
                                            want_history=False)

        for _ in range(args.warmup if warmup is None else warmup):
            step()
        synchronize()
        times, results = [], []
        for _ in range(args.steps if steps is None else steps):
            synchronize()
            t0 = time.perf_counter()
            r = step()
            synchronize()
            times.append((time.perf_counter() - t0) * 1e3)
            results.append(r)
        return float(np.median(times)), results, out.numpy()

    mp = size * size / 1e6
    if args.learned_only:
        from paintop_amd import warmstart
        warmstart.enable()
        ms, results, _ = timed(hip.PH_FLAG_WARM_START_NET)
        rl = results[len(results) // 2]
        print(json.dumps({"metric": METRIC + " (learned form only)", "value": round(ms / mp, 4),
                          "unit": "ms/MP", "n_gpus": 1, "steps": args.steps,
                          "warmup": args.warmup, "ms_per_step": round(ms, 3),
                          "config": {"workload": wl, "sweeps": rl.iterations,
                                     "launches": rl.launches,
                                     "coarse_corrections": rl.coarse_corrections,
                                     "warm_start_channels_kept": rl.warm_start_used}}),
              flush=True)
        return
    learned = warm_only = None
    cold_ms = None
    if args.workload == "seamless" and not args.plain_only:
        from paintop_amd import warmstart
        if os.path.exists(warmstart.WEIGHTS_FILE):
            warmstart.enable()
            learned = timed(hip.PH_FLAG_WARM_START_NET)
            # the learned start alone (no coarse corrections), for comparison
            os.environ["PAINTOP_HIP_CORRECTIONS"] = "0"
            try:
                warm_only = timed(hip.PH_FLAG_WARM_START_NET, steps=max(1, args.steps // 2),
                                  warmup=0)
            finally:
                os.environ.pop("PAINTOP_HIP_CORRECTIONS", None)
            # a mask the SOR-factor tuner has not seen: its coarse-mask eigen
            # solve runs inside the call (outside it, the cached value is used)
            os.environ["PAINTOP_HIP_TUNE_CACHE"] = "0"
            try:
                cold_ms = timed(hip.PH_FLAG_WARM_START_NET, steps=1, warmup=0)[0]
            finally:
                os.environ.pop("PAINTOP_HIP_TUNE_CACHE", None)
    ms_p, res_p, out_p = timed(0)
    r = res_p[len(res_p) // 2]
    ms, results = (learned[0], learned[1]) if learned else (ms_p, res_p)
    rl = results[len(results) // 2]
    # HBM traffic per sweep launch, measured by the rocprofv3 FETCH_SIZE/WRITE_SIZE
    # passes recorded in profiles/ (null when no measurement for this kernel)
    K = r.half_sweeps
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_sor_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        if (tr.get("kernel", "").startswith(f"k_sor<{K},") and
                tr.get("bench_workload", "seamless") == args.workload):
            traffic = tr["hbm_bytes_per_launch"]
    # roofline of the sweep kernel (plain form): each launch carries K/2 sweeps
    # per HBM round trip, so its compulsory traffic is 12 B per unknown per
    # launch (e read, e write, g read); timed by hipEvents around the sweep
    # launches (k_sor + its control kernel) on the solve's stream
    sweep_ms = float(np.median([x.sweep_ms for x in res_p]))
    launches = max(r.launches, 1)
    compulsory = BYTES_PER_UNKNOWN_SWEEP * r.unknown_sweeps / (K / 2)
    achieved = compulsory / (sweep_ms * 1e-3) / 1e9 if sweep_ms > 0 else 0.0
    effective = BYTES_PER_UNKNOWN_SWEEP * r.unknown_sweeps / (sweep_ms * 1e-3) / 1e9
    per_launch_us = sweep_ms * 1e3 / launches
    cfg = {
        "workload": wl,
        "form": "learned warm start + learned coarse corrections" if learned else
                "plain red-black SOR start",
        "sweeps": rl.iterations,
        "refinements": rl.refinements,
        "launches": rl.launches,
        "omega": round(rl.omega, 6),
        "half_sweeps_per_launch": rl.half_sweeps,
        "unknowns_per_channel": r.unknowns // (3 if chans == 4 else 1) if r.unknowns else 0,
        "converged": rl.converged,
        "final_residual": rl.final_residual,
        "device": info["name"],
    }
    if learned:
        rw = warm_only[1][len(warm_only[1]) // 2]
        cfg.update({
            "coarse_corrections": rl.coarse_corrections,
            "warm_start_channels_kept": rl.warm_start_used,
            "warm_start_initial_rel_residual": rl.warm_start_rel,
            "plain_ms_per_mp": round(ms_p / mp, 4),
            "plain_sweeps": r.iterations,
            "learned_speedup_vs_plain": round(ms_p / ms, 3),
            "learned_sweep_ratio": round(r.iterations / max(rl.iterations, 1), 3),
            "warm_start_only_sweeps": rw.iterations,
            "warm_start_only_ms_per_mp": round(warm_only[0] / mp, 4),
            "warm_start_only_speedup_vs_plain": round(ms_p / warm_only[0], 3),
            "cold_tuner_ms": round(cold_ms, 3),
            "cold_tuner_note": ("one learned solve with the SOR-factor tuner's cache "
                                "bypassed: the cost of the first solve on a new mask "
                                "(the workspace already allocated)"),
        })
    line = {
        "metric": METRIC,
        "value": round(ms / mp, 4),
        "unit": "ms/MP",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 sweeps, f64 residual/refinement" + (", bf16 network" if learned else ""),
        "data": "synthetic (device-generated fbm + star mask)",
        "config": cfg,
        "roofline": {
            "kernel": f"k_sor<{K},3> + k_sor_control<{K}> (one sweep launch, plain form)",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": traffic,
            "avg_launch_us": round(per_launch_us, 2),
            "bytes_per_launch_compulsory": round(compulsory / launches),
            "effective_GBps_12B_per_unknown_sweep": round(effective, 1),
            "measured_traffic_GBps": (round(traffic * launches / (sweep_ms * 1e-3) / 1e9, 1)
                                      if traffic and sweep_ms > 0 else None),
            "note": (f"achieved = compulsory bytes (12 B per unknown per launch: each launch "
                     f"runs {K // 2} sweeps per HBM round trip) / hipEvent time of the sweep "
                     f"launches; effective_GBps_12B_per_unknown_sweep charges 12 B to every "
                     f"unknown-sweep; measured_traffic_GBps = rocprofv3 FETCH+WRITE bytes per "
                     f"launch (profiles/k_sor_traffic.json) / the same time"),
        },
    }
    # error against the independent FP64 solution (the oracle's multigrid-CG
    # solve to 1e-12 on the CPU, sampled in tests/golden/)
    fx = {("seamless", 4096): "seamless_clone_4096_tight.npz",
          ("screened", 8192): "screened_8192_tight.npz"}.get((args.workload, size))
    if fx:
        z = np.load(os.path.join(ROOT, "tests", "golden", fx))
        pts = z["pts"]
        chan_tags = [k[:-6] for k in z.files if k.endswith("_u_pts")]

        def err_of(a):
            a = a.reshape(size * size, -1)
            return max(float(np.abs(a[pts, int(c[1:])].astype(np.float64) -
                                    z[f"{c}_u_pts"]).max()) for c in chan_tags)

        cfg["max_abs_err_vs_independent_fp64"] = err_of(out_p)
        if learned:
            cfg["learned_max_abs_err_vs_independent_fp64"] = err_of(learned[2])
        cfg["independent_fp64"] = (f"tests/golden/{fx}: oracle multigrid-preconditioned CG to "
                                   f"1e-12 (FP64 floor), 4096 sampled interior cells per channel")
    for a in (src, tgt, msk, out):
        a.free()
    if not args.no_cpu and args.workload == "seamless":
        line["cpu_baseline"] = cpu_baselines(size, r.omega, r.iterations, args.cpu_seconds)
        cb = line["cpu_baseline"]["value"]
        cfg["gpu_vs_cpu_optimized"] = round(cb / (ms / mp), 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
