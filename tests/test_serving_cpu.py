"""Scheduling logic of the serving loop, with stand-in solvers (no GPU)."""
import threading
import time

import numpy as np
import pytest

from at3d_amd.serving import Request, SolveServer


class FakeRTE:
    live = 0
    busy = 0
    busy_max = 0
    lock = threading.Lock()

    def __init__(self, n, fail=False, nbytes=1):
        self.n, self.fail, self.warm, self.nbytes = n, fail, None, nbytes
        self.num_iterations, self.solcrit = 0, 1.0
        with FakeRTE.lock:
            FakeRTE.live += 1

    def footprint_bytes(self):
        return self.nbytes

    def init_solution_from(self, other):
        self.warm = other.n

    def solve(self, maxiter, init_solution=True, verbose=False):
        with FakeRTE.lock:
            FakeRTE.busy += 1
            FakeRTE.busy_max = max(FakeRTE.busy_max, FakeRTE.busy)
        time.sleep(0.01)
        with FakeRTE.lock:
            FakeRTE.busy -= 1
        if self.fail:
            raise RuntimeError("boom")
        self.num_iterations = 3 if self.warm is not None else 7
        self.solcrit = 1e-5

    def integrate_to_sensor(self, sensor):
        return {"I": np.full(2, float(self.n))}

    def release(self):
        with FakeRTE.lock:
            FakeRTE.live -= 1


def test_server_runs_all_requests_with_warm_restarts_and_frees_solutions():
    FakeRTE.live = 0
    reqs = [Request(0, lambda: FakeRTE(0)), Request(1, lambda: FakeRTE(1), warm_from=0),
            Request(2, lambda: FakeRTE(2), warm_from=0, render={}),
            Request(3, lambda: FakeRTE(3), arrival_s=0.02), Request(4, lambda: FakeRTE(4), warm_from=3, arrival_s=0.03)]
    res, st = SolveServer(nworkers=3).run(reqs)
    assert [r.key for r in res] == [0, 1, 2, 3, 4]
    assert all(r.error is None for r in res)
    assert [r.iterations for r in res] == [7, 3, 3, 7, 3]
    assert np.array_equal(res[2].radiance, [2.0, 2.0])
    assert st["solves"] == 5 and st["failed"] == 0
    assert st["latency_p99_ms"] >= st["latency_p50_ms"] > 0
    assert FakeRTE.live == 0          # every device state released


def test_server_reports_failures_per_request():
    FakeRTE.live = 0
    reqs = [Request("a", lambda: FakeRTE(0, fail=True)), Request("b", lambda: FakeRTE(1)),
            Request("c", lambda: FakeRTE(2), warm_from="a")]
    res, st = SolveServer(nworkers=2).run(reqs)
    assert "boom" in res[0].error
    assert res[1].error is None
    assert "no solution" in res[2].error
    assert st["failed"] == 2
    assert FakeRTE.live == 0


def test_server_rejects_warm_start_from_a_later_request():
    with pytest.raises(ValueError):
        SolveServer().run([Request(0, lambda: FakeRTE(0), warm_from=1),
                           Request(1, lambda: FakeRTE(1), arrival_s=1.0)])


def test_server_admits_by_memory_footprint():
    """With a budget of 100 bytes, solves of 60 bytes run one at a time, 40-byte
    ones two at a time, and a 150-byte solve (over budget) still runs, alone; a
    kept warm-start source counts against the budget until its last user
    starts."""
    FakeRTE.live = FakeRTE.busy = FakeRTE.busy_max = 0
    reqs = [Request(k, (lambda k=k: FakeRTE(k, nbytes=60))) for k in range(4)]
    res, st = SolveServer(nworkers=4, memory_budget=100).run(reqs)
    assert all(r.error is None for r in res)
    assert FakeRTE.busy_max == 1 and st["peak_device_bytes"] == 60
    FakeRTE.busy_max = 0
    reqs = [Request(k, (lambda k=k: FakeRTE(k, nbytes=40))) for k in range(6)]
    res, st = SolveServer(nworkers=4, memory_budget=100).run(reqs)
    assert FakeRTE.busy_max == 2 and st["peak_device_bytes"] <= 100
    FakeRTE.busy_max = 0
    reqs = [Request("big", lambda: FakeRTE(0, nbytes=150)), Request("small", lambda: FakeRTE(1, nbytes=10))]
    res, st = SolveServer(nworkers=2, memory_budget=100).run(reqs)
    assert all(r.error is None for r in res) and FakeRTE.busy_max == 1
    reqs = [Request(0, lambda: FakeRTE(0, nbytes=50)), Request(1, lambda: FakeRTE(1, nbytes=50), warm_from=0),
            Request(2, lambda: FakeRTE(2, nbytes=50), warm_from=0)]
    res, st = SolveServer(nworkers=3, memory_budget=100).run(reqs)
    assert all(r.error is None for r in res) and [r.iterations for r in res] == [7, 3, 3]
    assert st["peak_device_bytes"] <= 100
    assert FakeRTE.live == 0


def test_group_latency_is_the_last_solve_of_the_group():
    reqs = [Request((0, j), lambda: FakeRTE(0), group=0) for j in range(3)] + \
           [Request((1, 0), lambda: FakeRTE(1), group=1, arrival_s=0.05)]
    res, st = SolveServer(nworkers=1).run(reqs)
    assert st["requests"] == 2 and st["solves"] == 4
    assert st["latency_p99_ms"] >= 25.0     # group 0 waits for three serial solves of 10 ms


def test_stream_requests_name_their_geometry_and_warm_builds_one_plan_each():
    """serving_stream tags every request's factory with (grid, NMU, NPHI);
    warm_geometries builds one plan per distinct geometry (host only)."""
    from at3d_amd import synthetic as S
    reqs = S.serving_stream(6, rate_hz=100.0, seed=2, grids=((8, 8, 6), (10, 8, 6)), angles=((4, 8),),
                            max_wavelengths=2)
    geoms = {r.make_rte.geometry for r in reqs}
    assert geoms <= {((8, 8, 6), 4, 8), ((10, 8, 6), 4, 8)} and len(geoms) >= 1
    # warm restarts share their source's geometry
    assert all(r.make_rte.geometry in geoms for r in reqs)
    assert S.warm_geometries(reqs) == len(geoms)


def test_failed_dependent_releases_its_source():
    """A warm-start request that fails before starting from its source (here
    its host setup raises) still counts as started: the kept solution is
    released once no dependent is left, and its bytes leave the budget."""
    FakeRTE.live = 0

    def bad():
        raise ValueError("bad medium")

    reqs = [Request("a", lambda: FakeRTE(0, nbytes=10)), Request("b", bad, warm_from="a"),
            Request("c", lambda: FakeRTE(2, nbytes=10), warm_from="a", arrival_s=0.05),
            Request("d", lambda: FakeRTE(3, nbytes=10), arrival_s=0.1)]
    res, st = SolveServer(nworkers=1, memory_budget=25).run(reqs)
    assert [r.error is None for r in res] == [True, False, True, True]
    assert res[2].iterations == 3
    assert FakeRTE.live == 0
    # "a" is released when "c" starts from it, so "d" never waits on it
    assert st["peak_device_bytes"] <= 20


def test_source_released_when_every_dependent_fails():
    FakeRTE.live = 0

    def bad():
        raise ValueError("bad medium")

    reqs = [Request("a", lambda: FakeRTE(0, nbytes=10)), Request("b", bad, warm_from="a"),
            Request("c", bad, warm_from="a"),
            Request("d", lambda: FakeRTE(3, nbytes=20), arrival_s=0.1)]
    res, st = SolveServer(nworkers=1, memory_budget=25).run(reqs)
    assert [r.error is None for r in res] == [True, False, False, True]
    assert FakeRTE.live == 0
    assert st["peak_device_bytes"] <= 20
