"""Serving many solves on one MI355X: continuous batching over device workers.

A solve of one medium leaves most of the GPU idle during its sweeps (one
workgroup per ordinate, a few hundred workgroups).  ``SolveServer`` keeps
``nworkers`` solves in flight, each on its own device state and HIP stream
(driven from its own host thread; the library releases the interpreter lock
in every call), and hands a worker the next request as soon as it finishes:
continuous batching.  A request may name an earlier request whose solution
seeds it (warm restart, ``RTE.init_solution_from``), as the reference's
optimisation loops continue from the last solution; a solution is kept on the
device only until every request that starts from it has started.  Requests
carry arrival times (for example a Poisson process); latency is measured from
arrival, service time from the moment a worker takes the request.

This module has no counterpart in the reference, whose solves run one per
thread on the CPU (at3d/containers.py:699-744).
"""
from __future__ import annotations

import queue
import threading
import time
from collections import Counter
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Request:
    key: object                    # caller's identifier (unique)
    make_rte: object               # () -> RTE
    maxiter: int = 100
    warm_from: object = None       # key of an earlier request whose solution seeds this one
    render: object = None          # optional sensor dict rendered after the solve
    arrival_s: float = 0.0         # arrival time relative to the start of run()
    group: object = None           # requests of one group (e.g. the wavelengths of one
                                   # observation) complete together: latency per group


@dataclass
class Result:
    key: object
    iterations: int
    solcrit: float
    latency_s: float               # completion - arrival
    service_s: float               # completion - start of service
    setup_s: float = 0.0           # host setup (make_rte) within the service time
    radiance: np.ndarray | None = None
    error: str | None = None


class SolveServer:
    """``nworkers`` host threads, each with its own device states and HIP
    streams; a request is admitted when a worker is free and its device
    footprint (``RTE.footprint_bytes``) fits in ``memory_budget`` bytes beside
    the solves in flight and the solutions kept for warm restarts (a request
    is admitted regardless when nothing else is in flight, so one larger than
    the budget still runs, alone).  ``memory_budget`` None: 80 % of the
    device memory free at the start of ``run``."""

    def __init__(self, nworkers=4, memory_budget=None):
        self.nworkers = int(nworkers)
        self.memory_budget = memory_budget

    def run(self, requests):
        """Serve ``requests``; returns (results in request order, stats)."""
        reqs = sorted(requests, key=lambda r: r.arrival_s)
        keys = [r.key for r in requests]
        if len(set(keys)) != len(keys):
            raise ValueError("request keys must be unique")
        order = {r.key: i for i, r in enumerate(reqs)}
        for r in reqs:
            if r.warm_from is not None and order.get(r.warm_from, len(reqs)) >= order[r.key]:
                raise ValueError(f"request {r.key!r} warm-starts from a request that does not arrive before it")
        pending = Counter(r.warm_from for r in reqs if r.warm_from is not None)
        q: queue.Queue = queue.Queue()
        results: dict = {}
        kept: dict = {}
        held: dict = {}      # key -> bytes of device memory it holds (running or kept)
        running = [0]
        peak = [0]
        budget = self.memory_budget
        if budget is None:
            budget = float("inf")
            try:
                import torch
                if torch.cuda.is_available():
                    budget = int(0.8 * torch.cuda.mem_get_info()[0])
            except Exception:
                pass
        cond = threading.Condition()
        t0 = time.perf_counter()

        def release_source(key):
            # one dependent of `key` has started (or failed): the kept
            # solution goes once no dependent is left.  Caller holds `cond`.
            pending[key] -= 1
            if pending[key] <= 0 and key in kept:
                kept.pop(key).release()
                held.pop(key, None)
                cond.notify_all()

        def admit(key, nbytes):
            with cond:
                while running[0] > 0 and sum(held.values()) + nbytes > budget:
                    cond.wait()
                held[key] = nbytes
                running[0] += 1
                peak[0] = max(peak[0], sum(held.values()))

        def worker():
            while True:
                req = q.get()
                if req is None:
                    return
                ts = time.perf_counter()
                rte = None
                setup = 0.0
                admitted = False
                source_done = req.warm_from is None
                try:
                    rte = req.make_rte()           # host setup: no device memory yet
                    setup = time.perf_counter() - ts
                    src = None
                    if req.warm_from is not None:
                        # wait for the source before taking device memory, so a
                        # waiting request never holds memory its source needs
                        with cond:
                            while req.warm_from not in results:
                                cond.wait()
                            src = kept.get(req.warm_from)
                        if src is None:
                            raise RuntimeError(f"request {req.warm_from!r} has no solution to start from")
                    fp = getattr(rte, "footprint_bytes", None)
                    admit(req.key, fp() if fp is not None else 0)
                    admitted = True
                    if src is not None:
                        try:
                            rte.init_solution_from(src)
                        finally:
                            with cond:
                                source_done = True
                                release_source(req.warm_from)
                        rte.solve(req.maxiter, init_solution=False, verbose=False)
                    else:
                        rte.solve(req.maxiter, verbose=False)
                    rad = None
                    if req.render is not None:
                        rad = rte.integrate_to_sensor(dict(req.render))["I"]
                    tn = time.perf_counter()
                    res = Result(req.key, rte.num_iterations, rte.solcrit,
                                 tn - (t0 + req.arrival_s), tn - ts, setup, rad, None)
                except Exception as e:  # reported per request
                    tn = time.perf_counter()
                    res = Result(req.key, 0, float("nan"), tn - (t0 + req.arrival_s), tn - ts,
                                 0.0, None, f"{type(e).__name__}: {e}")
                with cond:
                    if not source_done:
                        # failed before it could start from its source
                        release_source(req.warm_from)
                    if admitted:
                        running[0] -= 1
                    if res.error is None and pending.get(req.key, 0) > 0:
                        kept[req.key] = rte
                    else:
                        if rte is not None:
                            rte.release()
                        held.pop(req.key, None)
                    results[req.key] = res
                    cond.notify_all()

        threads = [threading.Thread(target=worker, daemon=True) for _ in range(self.nworkers)]
        for t in threads:
            t.start()
        for r in reqs:
            delay = t0 + r.arrival_s - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            q.put(r)
        for _ in threads:
            q.put(None)
        for t in threads:
            t.join()
        wall = time.perf_counter() - t0
        for rte in kept.values():
            rte.release()
        res = [results[k] for k in keys]
        ok = [r for r in res if r.error is None]
        # latency per group: the last of its solves to complete
        glat: dict = {}
        for rq in reqs:
            r = results[rq.key]
            if r.error is None:
                g = rq.key if rq.group is None else ("g", rq.group)
                glat[g] = max(glat.get(g, 0.0), r.latency_s)
        lat = np.array(list(glat.values()))
        svc = np.array([r.service_s for r in ok])
        stats = {"solves": len(res), "failed": len(res) - len(ok), "wall_s": wall,
                 "requests": len(glat), "peak_device_bytes": peak[0], "memory_budget": budget,
                 "solves_per_s": len(ok) / wall,
                 "latency_p50_ms": float(np.percentile(lat, 50) * 1e3) if lat.size else None,
                 "latency_p99_ms": float(np.percentile(lat, 99) * 1e3) if lat.size else None,
                 "setup_mean_ms": float(np.mean([r.setup_s for r in ok]) * 1e3) if ok else None,
                 "service_mean_ms": float(svc.mean() * 1e3) if svc.size else None,
                 "service_p50_ms": float(np.percentile(svc, 50) * 1e3) if svc.size else None,
                 "service_p99_ms": float(np.percentile(svc, 99) * 1e3) if svc.size else None}
        return res, stats
