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


def cpu_description():
    """Host CPU model, and the reference build's compiler and flags."""
    import platform
    import subprocess
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    comp = "amdflang"
    try:
        out = subprocess.run(["/opt/rocm/bin/amdflang", "--version"], capture_output=True, text=True,
                             timeout=20).stdout.splitlines()
        comp = out[0].strip() if out else comp
    except Exception:
        pass
    return {"cpu": model, "compiler": comp, "flags": "-O2 -fPIC (oracle/Makefile)"}


def cpu_baseline(workload_name: str, runs: int):
    """The reference Fortran (oracle/_ref) on one host core: setup,
    INIT_SOLUTION and SOLUTION_ITERATIONS to its own convergence on the same
    workload, `runs` times; the median run is the baseline.  Returns the
    baseline and the last run's converged state (for the render leg)."""
    from oracle import ref_shdom as R
    from at3d_amd import synthetic as S
    if not R.available() or runs <= 0:
        return None, None
    w = S.workload(workload_name)
    times, st = [], None
    for _ in range(runs):
        t0 = time.perf_counter()
        st = R.setup(R.scene_from_workload(w))
        R.init_solution(st)
        R.solution_iterations(st, w.maxiter)
        times.append(time.perf_counter() - t0)
        print(f"bench: reference solve {len(times)}/{runs}: {times[-1]:.1f} s, {st.iters} iterations",
              file=sys.stderr, flush=True)
    med = float(np.median(times))
    out = {"value": 1.0 / med, "unit": "solves/s", "cores": 1, "kind": "reference",
           "sample": f"reference Fortran on one host core: setup + INIT_SOLUTION + SOLUTION_ITERATIONS "
                     f"to its own convergence ({st.iters} iterations, SOLCRIT {st.solcrit:.2e}) of the "
                     f"same workload, median of {runs} runs",
           "runs_s": times, "iterations": int(st.iters)}
    out.update(cpu_description())
    return out, st


def cpu_render(st, rays):
    """The reference RENDER of the same rays from its converged solution, one host core."""
    from oracle import ref_shdom as R
    t0 = time.perf_counter()
    R.render(st, rays["ray_x"], rays["ray_y"], rays["ray_z"], rays["ray_mu"], rays["ray_phi"])
    return time.perf_counter() - t0


def gradient_case():
    """The differentiation workload: the 16^3 lognormal blob (NMU 8 / NPHI 16,
    64-term HG(0.85)), 3 views of 32 x 32 pixels, measurements rendered from
    the same generator with seed 1; unknowns extinction and albedo."""
    from at3d_amd import synthetic as S
    w = S.workload("blob16")
    target = S.lognormal_blob(n=16, nleg=64, seed=1)
    rays = S.workload_views("blob16", w)
    t = S.make_rte(target)
    t.solve(target.maxiter, verbose=False)
    meas = t.integrate_to_sensor(dict(rays))["I"].astype(np.float32)
    t.release()
    return w, rays, meas


def cpu_gradient(w, rays, meas):
    """The reference LEVISAPPROX_GRADIENT of the differentiation workload on one
    host core (its solve untimed)."""
    from oracle import ref_shdom as R
    if not R.available():
        return None
    st = R.setup(R.scene_from_workload(w))
    R.init_solution(st)
    R.solution_iterations(st, w.maxiter)
    shape = w.shape
    nleg = w.species[0][2].shape[0]
    R.prepare_derivatives(st, [
        R.Derivative(0, np.ones(shape), np.zeros(shape), np.zeros((nleg, 1)), np.ones(shape, np.int32)),
        R.Derivative(0, np.zeros(shape), np.ones(shape), np.zeros((nleg, 1)), np.ones(shape, np.int32))])
    t0 = time.perf_counter()
    R.levis_gradient(st, rays["ray_x"], rays["ray_y"], rays["ray_z"], rays["ray_mu"], rays["ray_phi"], meas)
    return time.perf_counter() - t0


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


def split_leg(cpu_runs):
    """Adaptive cell splitting (split_accuracy 0.03, the reference's default)
    on a 32 x 32 x 20 open cloud at L = 15: one solve of a new medium timed
    whole on the GPU (plan, state, iterations, every refinement of the grid),
    and, when the CPU baseline runs, the reference on one host core."""
    from at3d_amd import synthetic as S
    import torch
    w = S.cloud_field(nx=32, ny=32, nz=20, dx=0.04, dz=0.05, z0=0.3, seed=3, ext_scale=20.0,
                      solacc=1e-4)
    S.make_rte(w, split_accuracy=0.03).solve(w.maxiter, verbose=False)    # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rte = S.make_rte(w, split_accuracy=0.03)
    rte.solve(w.maxiter, verbose=False)
    torch.cuda.synchronize()
    out = {"workload": "32x32x20 open cloud, NMU16 NPHI32 L15, split_accuracy 0.03, solacc 1e-4",
           "base_points": rte.plan_size("npts"), "points": rte.state_size("npts"),
           "iterations": rte.num_iterations, "gpu_s": time.perf_counter() - t0}
    rte.release()
    if cpu_runs > 0:
        try:
            from oracle import ref_shdom as ref
        except Exception:
            return out
        if ref.available():
            t0 = time.perf_counter()
            st = ref.setup(ref.scene_from_workload(w, splitacc=0.03))
            ref.init_solution(st)
            ref.solution_iterations(st, w.maxiter)
            out.update(cpu_reference_s=time.perf_counter() - t0, reference_points=st.npts,
                       reference_iterations=st.iters)
            out["speedup_vs_cpu"] = out["cpu_reference_s"] / out["gpu_s"]
    return out


POL_WAVELENGTHS = (0.47, 0.66, 0.86)


def polarized_leg(cpu_runs):
    """The polarized workload: a 128 x 128 x 64 open cloud (dx = dy = dz =
    0.02 km; six-element cloud phase matrix plus Rayleigh with depolarization
    0.03), NSTOKES 3, NMU 16 NPHI 32 (L = 15), solacc 1e-4, at 0.47, 0.66 and
    0.86 um.  The three wavelengths are solved together (one host thread and
    HIP stream each), timed whole from the plan to convergence; then 9 views
    of 128 x 128 pixels are rendered for I, Q, U at each wavelength.  The
    reference's polarized solve is timed on one host core on a bounded
    sample: the 20 x 20 x 64 corner column block of the 0.66 um medium (same
    angles and layers), to its own convergence."""
    from concurrent.futures import ThreadPoolExecutor
    from at3d_amd import synthetic as S
    import torch
    kw = dict(max_total_mb=60000.0, adapt_grid_factor=1.0)   # a fixed grid: no room for splitting
    ws = [S.polarized_cloud_field(wavelength=wl) for wl in POL_WAVELENGTHS]

    def solve(w):
        r = S.make_rte(w, **kw)
        r.solve(w.maxiter, verbose=False)
        return r
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with ThreadPoolExecutor(len(ws)) as ex:
        rtes = list(ex.map(solve, ws))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    rays = S.workload_views("polarized128", ws[0])
    t0 = time.perf_counter()
    outs = [r.integrate_to_sensor(dict(rays)) for r in rtes]
    torch.cuda.synchronize()
    render_s = time.perf_counter() - t0
    out = {"workload": "128x128x64 open cloud, NSTOKES 3, NMU16 NPHI32 L15, solacc 1e-4, "
                       "3 wavelengths (0.47 0.66 0.86 um), 9 views of 128x128 I/Q/U",
           "npts": rtes[0].plan_size("npts"),
           "sweep": ("level-ordered, ring of radiances ({:.0f} KB of LDS)" if rtes[0].timing("sweep_ring4") else
                     "level-ordered ({:.0f} KB of LDS)").format(rtes[0].timing("sweep_ring_lds") / 1024)
           if rtes[0].timing("sweep_ring") else "point-ordered",
           "iterations": [r.num_iterations for r in rtes], "solve_wall_s": wall,
           "solves_per_s": len(ws) / wall, "render_wall_s": render_s,
           "rays": int(rays["ray_x"].size) * len(ws),
           "device_GB_per_state": rtes[0].timing("device_bytes") / 1e9,

           "mean_QoverI": float(np.mean([np.abs(o["Q"]).mean() / o["I"].mean() for o in outs]))}
    for r in rtes:
        r.release()
    # kernel times of one more solve with its iterations launched one by one
    # (graph-replayed iterations carry no timing events)
    r = S.make_rte(ws[1], **kw)
    r.set_option("iteration_graph", 0)
    r.solve(ws[1].maxiter, verbose=False)
    out["kernel_ms_per_iteration"] = {k: r.timing(k) / max(r.num_iterations, 1)
                                      for k in ("sh_to_do", "sweep", "sweep_kernel", "do_to_sh", "source", "total")}
    # the sweep against the HBM roof: DO source and extinction read, radiance
    # written, 12 B per point, ordinate and Stokes component, per iteration
    sweep_bytes = float(r.plan_size("npts")) * r.plan_size("nang") * 3 * 12
    sk = out["kernel_ms_per_iteration"]["sweep_kernel"]
    if sk > 0:
        ach = sweep_bytes / (sk * 1e-3) / 1e9
        out["sweep_roofline"] = {"bound": "hbm", "achieved": ach, "peak": 8000.0, "unit": "GB/s",
                                 "frac": ach / 8000.0, "bytes_per_iteration": sweep_bytes}
    r.release()
    if cpu_runs > 0:
        try:
            from oracle import ref_shdom as ref
        except Exception:
            return out
        if ref.available():
            w = ws[1]
            sub = [(e[:20, :20].copy(), a[:20, :20].copy(), t) for (e, a, t) in w.species]
            ws_ = S.Workload("polarized_sample", w.delx, w.dely, w.z, sub, bc="open", nmu=w.nmu, nphi=w.nphi,
                             solacc=w.solacc, wavelength=w.wavelength, nstokes=3)
            t0 = time.perf_counter()
            st = ref.solve(ref.scene_from_workload(ws_), maxiter=w.maxiter)
            cpu_s = time.perf_counter() - t0
            per_pt = cpu_s / st.npts
            out["cpu_reference"] = {"sample": "20x20x64 corner block of the 0.66 um medium, to convergence",
                                    "sample_points": st.npts, "sample_s": cpu_s,
                                    "sample_iterations": st.iters, "cores": 1,
                                    "est_s_per_solve_full_grid": per_pt * out["npts"],
                                    "speedup_vs_cpu_est": per_pt * out["npts"] / (wall / len(ws))}
    return out


def cold_solves(n, maxiter):
    """`n` solves of new media on the headline geometry (the cloud field
    re-drawn with seeds 1..n): host plan of the new medium (the sweep levels
    of the geometry are cached), device state, records pass, initialisation
    and iterations to convergence, each timed whole."""
    from at3d_amd import synthetic as S
    import torch
    out = []
    for k in range(1, n + 1):
        wk = S.cloud_field(z0=0.35, seed=k)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = S.make_rte(wk)
        t1 = time.perf_counter()
        r.solve(maxiter, verbose=False)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        out.append({"plan_ms": (t1 - t0) * 1e3, "state_and_solve_ms": (t2 - t1) * 1e3,
                    "total_ms": (t2 - t0) * 1e3, "iterations": r.num_iterations,
                    "device_ms": r.timing("total")})
        r.release()
    return out


def render_leg(rte, w, reps):
    """The headline workload's 9 views of 64 x 64 pixels rendered from its
    converged solution `reps` times (after one untimed call)."""
    from at3d_amd import synthetic as S
    import torch
    rays = S.workload_views("cloud64", w)
    rte.integrate_to_sensor(dict(rays))
    torch.cuda.synchronize()
    wall, dev, kr = [], [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        rte.integrate_to_sensor(dict(rays))
        wall.append((time.perf_counter() - t0) * 1e3)
        dev.append(rte.timing("render"))
        kr.append(rte.timing("render_rays"))
    nrays = rays["ray_x"].size
    npts, nlm = rte._npts, rte._nlm
    ndir = 9
    # algorithmic bytes: the direction source reads each point's source terms
    # once and writes one value per (point, view); the rays read each point's
    # extinction and direction source once per view and read / write their
    # own 36 bytes (x, y, z, mu, phi, radiance)
    ns = rte.state_array("ns").astype(np.float64)
    b_dir = 4.0 * ns.sum() + 4.0 * ndir * npts
    b_rays = 8.0 * ndir * npts + 36.0 * nrays
    ms_rays = float(np.median(kr))
    return rays, {"views": ndir, "rays": nrays, "wall_ms": float(np.median(wall)),
                  "device_ms": float(np.median(dev)), "ray_kernel_ms": ms_rays,
                  "rays_per_s": nrays / (float(np.median(wall)) * 1e-3),
                  "ray_kernel_bytes": b_rays, "ray_kernel_GB/s": b_rays / ms_rays / 1e6,
                  "ray_kernel_frac_hbm": b_rays / ms_rays / 1e6 / HBM_PEAK_GBS,
                  "direction_source_bytes": b_dir}


def gradient_leg(w, rays, meas, reps):
    """The Levis gradient of the differentiation workload (solve untimed),
    `reps` times after one untimed call."""
    from at3d_amd import gradient as GR
    from at3d_amd import synthetic as S
    rte = S.make_rte(w)
    rte.solve(w.maxiter, verbose=False)
    sp = list(rte.medium)[0]
    unknowns = [(sp, "extinction"), (sp, "ssalb")]
    sensor = dict(rays, measurement_data=meas)
    GR.levis_approximation_grad(rte, sensor, unknowns)
    wall, dev, adj = [], [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        GR.levis_approximation_grad(rte, sensor, unknowns)
        wall.append((time.perf_counter() - t0) * 1e3)
        dev.append(rte.timing("gradient"))
        adj.append(rte.timing("adjoint_rays"))
    rte.release()
    return {"pixels": int(meas.size), "unknowns": 2, "wall_ms": float(np.median(wall)),
            "device_ms": float(np.median(dev)), "adjoint_ray_kernel_ms": float(np.median(adj))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cloud64")
    ap.add_argument("--cpu-runs", type=int, default=3,
                    help="reference solves timed for cpu_baseline (the median is reported; 0: skip)")
    ap.add_argument("--cold-steps", type=int, default=3, help="solves of new media timed whole")
    ap.add_argument("--render-reps", type=int, default=3)
    ap.add_argument("--gradient-reps", type=int, default=3)
    ap.add_argument("--polarized", type=int, default=1, help="run the polarized (NSTOKES 3) leg")
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--stream-requests", type=int, default=512,
                    help="requests of the mixed-size serving stream (0: skip it; 512 is the "
                         "stream's stated size)")
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
    print("bench: warm-up done", file=sys.stderr, flush=True)

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
        iters_total += int(rte.timing("iters"))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_per_step = elapsed * 1e3 / args.steps
    iters = iters_total // args.steps
    graphs = bool(rte.timing("iteration_graph_active"))
    print("bench: plain solves done", file=sys.stderr, flush=True)
    # per-kernel times: the same solves with the iterations launched one by
    # one, HIP events on the state's stream between the kernel classes and
    # around each sweep kernel launch (the kernels are those the graphs run)
    rte.set_option("iteration_graph", 0)
    kiters = 0
    for _ in range(args.steps):
        rte.solve(w.maxiter, verbose=False)
        for k in keys:
            ksum[k] += rte.timing(k)
        kiters += int(rte.timing("iters"))
    rte.set_option("iteration_graph", 1)
    torch.cuda.synchronize()

    # engine variants of the same plain solve (for the design record): the
    # iterations launched one by one with the host deciding after each
    variants = {}
    for vname, opts, reset in (("host_controlled_iterations", {"iteration_graph": 0}, {"iteration_graph": 1}),):
        for k, v in opts.items():
            rte.set_option(k, v)
        rte.solve(w.maxiter, verbose=False)
        torch.cuda.synchronize()
        tv = time.perf_counter()
        for _ in range(args.steps):
            rte.solve(w.maxiter, verbose=False)
        torch.cuda.synchronize()
        variants[vname] = {"ms_per_solve": (time.perf_counter() - tv) * 1e3 / args.steps,
                           "iterations": rte.num_iterations}
        for k, v in reset.items():
            rte.set_option(k, v)
    print("bench: variants done", file=sys.stderr, flush=True)

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
    for r in servers[1:]:
        r.release()
    print("bench: served solves done", file=sys.stderr, flush=True)

    npts, nang, nlm = rte._npts, rte._nang, rte._nlm
    ns = rte.state_array("ns").astype(np.float64)
    nr = rte.state_array("nr").astype(np.float64)
    kern = {}
    n_it = max(kiters, 1)
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
    # the transforms on the terms that can be nonzero (each point's source
    # truncation NS for SH->DO, radiance truncation NR for DO->SH): useful
    # flops 2 NANG sum(NS) and 2 NANG sum(NR), bytes one read of those terms,
    # one read or write of NANG values per point
    for k, cnt in (("sh_to_do", ns), ("do_to_sh", nr)):
        ms = ksum[k] / n_it
        fl = 2.0 * nang * cnt.sum()
        by = 4.0 * (cnt.sum() + nang * npts)
        kern[k] = {"ms_per_launch": ms, "useful_flops_per_launch": fl,
                   "useful_TFLOP/s": fl / ms / 1e9, "bytes_per_launch": by, "GB/s": by / ms / 1e6}
    # the source pass reads R (NR terms), the old S (NS) and DS, writes S and
    # DS, and reads each point's Legendre terms and scalars
    ms_src = ksum["source"] / n_it
    src_bytes = 4.0 * (nr.sum() + 4.0 * ns.sum()) + npts * 4.0 * (rte._ml + 1 + 8)
    kern["source"] = {"ms_per_launch": ms_src, "bytes_per_launch": src_bytes,
                      "GB/s": src_bytes / ms_src / 1e6}
    dominant = max(("sweep", "sh_to_do", "do_to_sh", "source"), key=lambda k: ksum[k])
    traffic = measured_traffic(dominant)
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

    cold = cold_solves(args.cold_steps, w.maxiter) if args.cold_steps > 0 else []
    print("bench: cold solves done", file=sys.stderr, flush=True)
    rays, render = render_leg(rte, w, args.render_reps)
    print("bench: render done", file=sys.stderr, flush=True)
    gw, grays, gmeas = gradient_case()
    grad = gradient_leg(gw, grays, gmeas, args.gradient_reps)
    print("bench: gradient done", file=sys.stderr, flush=True)

    split = split_leg(args.cpu_runs)
    print("bench: adaptive splitting done", file=sys.stderr, flush=True)
    polarized = polarized_leg(args.cpu_runs) if args.polarized else None
    print("bench: polarized done", file=sys.stderr, flush=True)

    stream = None
    if args.stream_requests > 0:
        stream = serving_stream_bench(args.stream_requests, args.stream_rate, args.stream_workers)
    print("bench: serving stream done", file=sys.stderr, flush=True)

    cpu, ref_st = cpu_baseline(args.workload, args.cpu_runs)
    if cpu is not None:
        render["cpu_reference_s"] = cpu_render(ref_st, rays)
        render["speedup_vs_cpu"] = render["cpu_reference_s"] * 1e3 / render["wall_ms"]
        g_cpu = cpu_gradient(gw, grays, gmeas)
        if g_cpu is not None:
            grad["cpu_reference_s"] = g_cpu
            grad["speedup_vs_cpu"] = g_cpu * 1e3 / grad["wall_ms"]

    plain_value = args.steps / elapsed
    cold_ms = float(np.median([c["total_ms"] for c in cold])) if cold else None
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
        "cold": {"solves": cold, "median_ms": cold_ms},
        "variants": variants,
        "kernel_ms_per_solve": {k: ksum[k] / args.steps for k in keys},
        "iterations_from_graphs": graphs,
        "kernels": kern,
        "roofline": roof,
        "solve_roofline": solve_roof,
        "render": render,
        "gradient": grad,
        "adaptive_splitting": split,
        "polarized": polarized,
        "serving_stream": stream,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": {"cold_solve": (1e3 / cold_ms / cpu["value"]) if (cpu and cold_ms) else None,
                           "plain_solve": (plain_value / cpu["value"]) if cpu else None},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
