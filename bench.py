"""Benchmark: converged SHDOM solves/s on one MI355X, served and plain.

Workload (the configuration the headline metric is quoted on): scalar solar
solve on a 64x64x40 open base grid (dx=dy=0.02 km, dz=0.025 km), NMU=16,
NPHI=32, L=M=15, solution accuracy 1e-4, a seeded -5/3 Gaussian-field cloud
(30/km x field^1.5, 40 % cover, 200-term phase function) plus a 0.01/km
Rayleigh species.  A step is one converged solve: initial radiance and source
(device) + source-function iterations until SOLCRIT <= 1e-4.  The host setup
(grid, optics, direct beam, two-stream start, sweep plan) and the upload to
HBM happen before the timed regions and are reported as ``setup_ms``.

plain: one solver, --steps solves back to back; served: the same --steps
solves spread over --workers solvers, each with its own HIP stream and host
thread, all in flight together.  ``value`` is the served rate (--steps
converged solves over the served wall time); the plain rate and the per-solve
latency percentiles of both modes are in the same line.

Prints one JSON line.  ``--gpus N`` with N > 1 exits: this engine runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
FP32_MFMA_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix


def cpu_baseline(workload_name: str, iters_gpu: int, niter: int):
    """Times the reference Fortran (oracle/_ref) on one host core on the same
    workload: setup, INIT_SOLUTION and `niter` SOLUTION_ITERATIONS (0: as many
    as the GPU solve took, i.e. one whole solve), scaled to one converged solve
    of `iters_gpu` iterations."""
    from oracle import ref_shdom as R
    from at3d_amd import synthetic as S
    if not R.available():
        return None
    w = S.workload(workload_name)
    t0 = time.perf_counter()
    st = R.setup(R.scene_from_workload(w))
    R.init_solution(st)
    t1 = time.perf_counter()
    R.solution_iterations(st, niter if niter > 0 else iters_gpu)
    t2 = time.perf_counter()
    per_iter = (t2 - t1) / max(st.iters, 1)
    solve_s = (t1 - t0) + per_iter * iters_gpu
    return {"value": 1.0 / solve_s, "unit": "solves/s", "cores": 1, "kind": "reference",
            "sample": f"reference Fortran (amdflang -O2) setup+INIT_SOLUTION {t1 - t0:.2f} s and "
                      f"{st.iters} SOLUTION_ITERATIONS at {per_iter:.2f} s each on one host core; "
                      f"one converged solve of the GPU's {iters_gpu} iterations = {solve_s:.1f} s"}


# kernel class -> device kernel name in the rocprofv3 counter summaries
KERNEL_NAMES = {"sweep": "k_sweep_ring", "sh_to_do": "k_sh_to_do_perm", "do_to_sh": "k_do_to_sh",
                "source": "k_compute_source"}


def measured_traffic(kernel_class: str):
    """HBM bytes per launch of the kernel from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled
    by the reading of a copy of known size), or None when there is none."""
    import re
    path = os.path.join(ROOT, "profiles", "pmc_traffic_cloud64.md")
    if not os.path.exists(path):
        return None
    name = KERNEL_NAMES.get(kernel_class)
    with open(path) as f:
        for line in f:
            m = re.match(r"^\|\s*(\w+)\s*\|\s*([0-9.eE+]+)\s*\|\s*([0-9.eE+]+)\s*\|\s*([0-9.eE+]+)\s*\|",
                         line)
            if m and m.group(1) == name:
                return float(m.group(4))
    return None


def serving_stream_bench(n_requests, rate_hz, workers):
    """The mixed-size serving stream (grids 32^3 .. 96x96x48, (NMU, NPHI) from
    (8,16) to (16,32), 1-4 wavelengths per request, 30 % warm restarts with 5 %
    extinction noise, solution accuracy 1e-5), Poisson arrivals: served by
    `workers` solvers with admission by device footprint, and the same
    requests one at a time.  Host setup (plan) is part of every request."""
    from at3d_amd import synthetic as S
    from at3d_amd.serving import SolveServer
    from at3d_amd.solver import device_trim
    reqs = S.serving_stream(n_requests, rate_hz=rate_hz, seed=1)
    # both runs start with the sweep levels of every geometry of the stream
    # cached on the host (as a running server has them) and without kept
    # device blocks
    S.warm_geometries(reqs)
    device_trim()
    res, st = SolveServer(nworkers=workers).run(reqs)
    served_iters = {r.key: (r.iterations, r.solcrit) for r in res}
    for r in reqs:
        r.arrival_s = 0.0
    device_trim()
    res1, st1 = SolveServer(nworkers=1).run(reqs)
    same = all(served_iters[r.key] == (r.iterations, r.solcrit) for r in res1)
    return {"requests": st["requests"], "solves": st["solves"], "failed": st["failed"] + st1["failed"],
            "offered_rate_hz": rate_hz, "workers": workers,
            "served_solves_per_s": st["solves_per_s"], "one_at_a_time_solves_per_s": st1["solves_per_s"],
            "speedup": st["solves_per_s"] / st1["solves_per_s"],
            "latency_ms": {"p50": st["latency_p50_ms"], "p99": st["latency_p99_ms"]},
            "one_at_a_time_per_solve_ms": {"host_setup": st1["setup_mean_ms"],
                                           "total": st1["service_mean_ms"]},
            "peak_device_GB": st["peak_device_bytes"] / 1e9,
            "same_iterations_and_solcrit_as_one_at_a_time": same}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cloud64")
    ap.add_argument("--cpu-iters", type=int, default=0,
                    help="reference iterations timed for cpu_baseline (0: the GPU solve's count)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--stream-requests", type=int, default=48,
                    help="requests of the mixed-size serving stream (0: skip it)")
    ap.add_argument("--stream-rate", type=float, default=40.0, help="its Poisson arrival rate (1/s)")
    ap.add_argument("--stream-workers", type=int, default=8)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first copy a buffer of this size once (the known-size kernel that "
                         "scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this SHDOM engine runs one solve per GPU and is measured on one GPU; "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import torch
    from at3d_amd import synthetic as S

    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    w = S.workload(args.workload)
    t0 = time.perf_counter()
    rte = S.make_rte(w)
    rte.solve(0, verbose=False)         # allocates the device state, uploads, initialises
    torch.cuda.synchronize()
    setup_ms = (time.perf_counter() - t0) * 1e3

    for _ in range(args.warmup):
        rte.solve(w.maxiter, verbose=False)
    torch.cuda.synchronize()

    keys = ["truncation", "sh_to_do", "sweep", "sweep_kernel", "do_to_sh", "source", "total"]
    ksum = {k: 0.0 for k in keys}
    iters_total = 0
    sweep_levels = int(rte.timing("sweep_levels"))
    lat = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        rte.solve(w.maxiter, verbose=False)
        lat.append((time.perf_counter() - ts) * 1e3)
        for k in keys:
            ksum[k] += rte.timing(k)
        iters_total += int(rte.timing("iters"))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_per_step = elapsed * 1e3 / args.steps
    iters = iters_total // args.steps

    # served: the same number of solves over several solvers in flight, one
    # stream and host thread each, each taking the next solve when it is free
    import threading
    servers = [rte] + [S.make_rte(w) for _ in range(max(args.workers, 1) - 1)]
    for r in servers[1:]:
        r.solve(w.maxiter, verbose=False)
    torch.cuda.synchronize()
    served_lat = []
    lock = threading.Lock()
    todo = [args.steps]

    def work(r):
        while True:
            with lock:
                if todo[0] <= 0:
                    return
                todo[0] -= 1
            ts = time.perf_counter()
            r.solve(w.maxiter, verbose=False)
            with lock:
                served_lat.append((time.perf_counter() - ts) * 1e3)

    torch.cuda.synchronize()
    ts0 = time.perf_counter()
    th = [threading.Thread(target=work, args=(r,)) for r in servers]
    for t in th:
        t.start()
    for t in th:
        t.join()
    torch.cuda.synchronize()
    served_elapsed = time.perf_counter() - ts0
    served_n = args.steps
    served_value = served_n / served_elapsed

    npts, nang, nlm = rte._npts, rte._nang, rte._nlm
    # algorithmic bytes / flops per launch of each kernel class
    gemm_flops = 2.0 * nang * npts * nlm
    kern = {}
    n_it = max(iters_total, 1)
    # one sweep pass = one launch per hemisphere (a workgroup per ordinate walking
    # the octant's dependency levels); algorithmic bytes per (ordinate, point):
    # read the DO source and the extinction, write the DO radiance (3 x 4 B)
    sweep_bytes = nang * npts * 12.0
    ms_pass = ksum["sweep"] / n_it
    # the sweep kernel alone (events around its two launches per pass); the
    # pass also holds the surface boundary kernel
    ms_launch = ksum["sweep_kernel"] / n_it / 2 if ksum["sweep_kernel"] > 0 else ms_pass / 2
    kern["sweep"] = {"ms_per_pass": ms_pass, "launches_per_pass": 2,
                     "dependency_levels": sweep_levels,
                     "ms_per_launch": ms_launch, "bytes_per_launch": sweep_bytes / 2,
                     "GB/s": sweep_bytes / 2 / ms_launch / 1e6,
                     "level_ordered": bool(rte.timing("sweep_ring")),
                     "lds_bytes": int(rte.timing("sweep_ring_lds")),
                     "multi_cell_points": int(rte.timing("sweep_general_points"))}
    # dense-equivalent rates: both transforms skip terms beyond each point's
    # SH truncation (exact zeros), so they execute fewer flops than 2 N M K
    for k in ("sh_to_do", "do_to_sh"):
        ms = ksum[k] / n_it
        kern[k] = {"ms_per_launch": ms, "dense_flops_per_launch": gemm_flops,
                   "dense_equiv_TFLOP/s": gemm_flops / ms / 1e9}
    ms_src = ksum["source"] / n_it
    src_bytes = npts * nlm * 4 * 6.0       # R read, S read/write, DS read/write, chi/ysun
    kern["source"] = {"ms_per_launch": ms_src, "bytes_per_launch": src_bytes,
                      "GB/s": src_bytes / ms_src / 1e6}
    dominant = max(("sweep", "sh_to_do", "do_to_sh", "source"), key=lambda k: ksum[k])
    traffic = measured_traffic(dominant)
    if dominant in ("sh_to_do", "do_to_sh"):
        ach = kern[dominant]["dense_equiv_TFLOP/s"]
        roof = {"kernel": dominant, "bound": "mfma", "achieved": ach, "peak": FP32_MFMA_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": ach / FP32_MFMA_PEAK_TFLOPS, "traffic": traffic}
    else:
        ach = kern[dominant]["GB/s"]
        roof = {"kernel": dominant, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": traffic}

    # whole-solve algorithmic traffic (minimum bytes of a dense FP32 layout per
    # iteration: source r/w with acceleration 5 NLM, SH->DO NLM + NANG, sweep
    # 3 NANG, DO->SH NANG + NLM; plus the setup's optics and the first
    # source), over the plain solve time
    b_iter = npts * 4.0 * (7 * nlm + 5 * nang)
    b_solve = iters * b_iter + npts * 4.0 * (2 * nlm + 16)
    solve_gbs = b_solve / (ms_per_step * 1e-3) / 1e9
    solve_roof = {"bytes_per_solve": b_solve, "bytes_per_iteration": b_iter, "GB/s": solve_gbs,
                  "frac": solve_gbs / HBM_PEAK_GBS}

    stream = None
    if args.stream_requests > 0:
        stream = serving_stream_bench(args.stream_requests, args.stream_rate, args.stream_workers)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.workload, iters, args.cpu_iters)

    plain_value = args.steps / elapsed
    value = served_value
    line = {
        "metric": "converged solves/s @64x64x40 L=15 (served vs plain); p50/p99 latency; %HBM roof",
        "value": value, "unit": "solves/s", "n_gpus": 1, "steps": served_n,
        "warmup": args.warmup, "ms_per_step": served_elapsed * 1e3 / served_n,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "cloud64: 64x64x40 open grid, NMU16 NPHI32 L15, solacc 1e-4, "
                               "two species (cloud + Rayleigh)",
                   "npts": npts, "nang": nang, "nlm": nlm, "iterations": iters,
                   "setup_ms": setup_ms, "solcrit": rte.solcrit},
        "served": {"solves_per_s": served_value, "workers": len(servers), "solves": served_n,
                   "latency_ms": {"p50": float(np.percentile(served_lat, 50)),
                                  "p99": float(np.percentile(served_lat, 99))}},
        "plain": {"solves_per_s": plain_value, "solves": args.steps, "ms_per_solve": ms_per_step,
                  "latency_ms": {"p50": float(np.percentile(lat, 50)),
                                 "p99": float(np.percentile(lat, 99))}},
        "kernel_ms_per_solve": {k: ksum[k] / args.steps for k in keys},
        "kernels": kern,
        "roofline": roof,
        "solve_roofline": solve_roof,
        "serving_stream": stream,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
