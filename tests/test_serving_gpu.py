"""Warm restarts and concurrent serving on the GPU."""
import numpy as np
import pytest

from at3d_amd import synthetic as S
from at3d_amd.serving import Request, SolveServer
from tests.scenes import blob8, cloud_small, ortho_rays, rel_l2

pytestmark = pytest.mark.gpu


def _perturbed(w, seed, scale=0.05):
    rng = np.random.default_rng(seed)
    sp = []
    for ext, alb, leg in w.species:
        noise = np.exp(scale * rng.standard_normal(ext.shape) - 0.5 * scale ** 2)
        sp.append(((ext * noise).astype(np.float32), alb, leg))
    return S.Workload(w.name, w.delx, w.dely, w.z, sp, bc=w.bc, nmu=w.nmu, nphi=w.nphi,
                      solacc=w.solacc)


def test_warm_restart_converges_to_cold_solution_in_fewer_iterations():
    w0 = cloud_small()
    w1 = _perturbed(w0, seed=3)
    a = S.make_rte(w0)
    a.solve(100, verbose=False)
    cold = S.make_rte(w1)
    cold.solve(100, verbose=False)
    warm = S.make_rte(w1)
    warm.init_solution_from(a)
    warm.solve(100, init_solution=False, verbose=False)
    assert warm.check_solved(verbose=False)
    assert warm.num_iterations < cold.num_iterations
    assert rel_l2(warm.state_array("source"), cold.state_array("source")) < 2e-4


def test_warm_restart_rejects_other_grids():
    a = S.make_rte(blob8("open"))
    a.solve(2, verbose=False)
    b = S.make_rte(cloud_small())
    with pytest.raises(Exception, match="same grid"):
        b.init_solution_from(a)


def test_warm_restart_rejects_another_sh_layout_of_equal_size():
    """NMU 16 / NPHI 1 (L = 15, M = 0) and NMU 4 / NPHI 8 (L = M = 3) both hold
    16 spherical-harmonic terms per point, in different (l, m) orders."""
    a = S.make_rte(blob8("open", nmu=16, nphi=1))
    a.solve(2, verbose=False)
    b = S.make_rte(blob8("open", nmu=4, nphi=8))
    assert a.plan_size("nlm") == b.plan_size("nlm") == 16
    with pytest.raises(Exception, match="same grid"):
        b.init_solution_from(a)


def test_concurrent_serving_is_bit_identical_to_serial_solves():
    base = [blob8("open"), blob8("periodic"), cloud_small()]
    ws = base + [_perturbed(base[2], seed=s) for s in (1, 2, 3)]
    views = [ortho_rays(w, [(0.0, 0.0), (45.0, 30.0)]) for w in ws]
    serial = []
    for w, v in zip(ws, views):
        r = S.make_rte(w)
        r.solve(100, verbose=False)
        serial.append(r.integrate_to_sensor(dict(v))["I"])
        r.release()
    reqs = [Request(i, (lambda w=w: S.make_rte(w)), render=views[i]) for i, w in enumerate(ws)]
    res, st = SolveServer(nworkers=4).run(reqs)
    assert st["failed"] == 0, [r.error for r in res]
    for r, ref in zip(res, serial):
        assert np.array_equal(r.radiance, ref)
    # warm-restarted chain: each perturbed scene starts from the previous solution
    reqs = [Request("c0", lambda: S.make_rte(ws[2]))] + [
        Request(f"c{k}", (lambda w=ws[2 + k]: S.make_rte(w)), warm_from=f"c{k - 1}", render=views[2 + k])
        for k in (1, 2, 3)]
    res, st = SolveServer(nworkers=2).run(reqs)
    assert st["failed"] == 0, [r.error for r in res]
    for k in (1, 2, 3):
        assert rel_l2(res[k].radiance, serial[2 + k]) < 2e-4


def test_serving_stream_results_are_the_same_alone_and_in_the_mix():
    """A small mixed stream (two grids, two angle sets, up to three
    wavelengths per request, warm restarts) served by three solvers with a
    memory budget that keeps some requests waiting: every solve's iteration
    count, SOLCRIT and rendered radiances equal those of the same requests run
    one at a time."""
    reqs = S.serving_stream(8, rate_hz=50.0, seed=5, grids=((16, 16, 16), (24, 24, 16)),
                            angles=((8, 16), (12, 24)), max_wavelengths=3, solacc=1e-4)
    view = {"ray_x": np.linspace(0.1, 0.7, 6), "ray_y": np.full(6, 0.4), "ray_z": np.full(6, 1.0),
            "ray_mu": np.full(6, 1.0), "ray_phi": np.zeros(6)}
    for r in reqs:
        r.render = view
    fps = [r.make_rte().footprint_bytes() for r in reqs]
    budget = 2 * max(fps)
    res, st = SolveServer(nworkers=3, memory_budget=budget).run(reqs)
    assert st["failed"] == 0, [r.error for r in res]
    # over the budget only by the one request admitted when nothing else runs
    assert st["peak_device_bytes"] <= budget + max(fps)
    for r in reqs:
        r.arrival_s = 0.0
    res1, st1 = SolveServer(nworkers=1).run(reqs)
    assert st1["failed"] == 0
    for a, b in zip(res, res1):
        assert (a.iterations, a.solcrit) == (b.iterations, b.solcrit)
        assert np.array_equal(a.radiance, b.radiance)


def test_footprint_estimate_covers_the_state():
    """The footprint known before allocation is within 10 % above the state's
    device allocations (it is what admission reserves)."""
    for w in (blob8("open"), cloud_small(), S.serve_medium((24, 24, 16), 12, 24, 1.0, 0)):
        r = S.make_rte(w)
        r.solve(0, verbose=False)
        est, got = r.footprint_bytes(), r.timing("device_bytes")
        assert got <= est <= 1.1 * got + (1 << 20), (est, got)
        r.release()


def test_stream_over_every_grid_and_angle_set_is_the_same_alone_and_in_the_mix():
    """The serving workload's shapes: each of the four grids (32^3, 48^2 x 32,
    64^2 x 40, 96^2 x 48) and each of the three angle sets ((NMU, NPHI, L) =
    (8, 16, 7), (12, 24, 11), (16, 32, 15)), one to three wavelengths per
    request, two warm restarts from perturbed earlier requests, served by
    four solvers under a memory budget that keeps requests waiting; every
    solve's iteration count, SOLCRIT and rendered radiances equal those of
    the same requests run one at a time."""
    specs = [((32, 32, 32), 8, 16, [0.86]), ((48, 48, 32), 12, 24, [0.55, 1.6]),
             ((64, 64, 40), 16, 32, [0.86]), ((96, 96, 48), 8, 16, [2.1]),
             ((32, 32, 32), 16, 32, [0.67, 1.2, 2.2]), ((96, 96, 48), 12, 24, [0.9]),
             ((48, 48, 32), 16, 32, [1.1])]
    warm = {7: 2, 8: 1}
    view = {"ray_x": np.linspace(0.05, 0.75, 8), "ray_y": np.full(8, 0.4), "ray_z": np.full(8, 0.64),
            "ray_mu": np.full(8, 0.9), "ray_phi": np.full(8, 0.5)}
    reqs = []
    for i in range(len(specs) + len(warm)):
        if i < len(specs):
            grid, nmu, nphi, wls = specs[i]
            seed, noise, src = 100 + i, None, None
        else:
            src = warm[i]
            grid, nmu, nphi, wls = specs[src]
            seed, noise = 100 + src, 0.05
        for j, wl in enumerate(wls):
            def make(grid=grid, nmu=nmu, nphi=nphi, wl=wl, seed=seed, noise=noise, i=i):
                return S.make_rte(S.serve_medium(grid, nmu, nphi, wl, seed, noise, noise_seed=i,
                                                 solacc=1e-4))
            reqs.append(Request((i, j), make, warm_from=None if src is None else (src, j),
                                render=view, arrival_s=0.01 * i, group=i))
    fps = [r.make_rte().footprint_bytes() for r in reqs]
    budget = 2 * max(fps)
    res, st = SolveServer(nworkers=4, memory_budget=budget).run(reqs)
    assert st["failed"] == 0, [r.error for r in res]
    assert st["peak_device_bytes"] <= budget + max(fps)
    for r in reqs:
        r.arrival_s = 0.0
    res1, st1 = SolveServer(nworkers=1).run(reqs)
    assert st1["failed"] == 0, [r.error for r in res1]
    for a, b in zip(res, res1):
        assert (a.iterations, a.solcrit) == (b.iterations, b.solcrit), a.key
        assert np.array_equal(a.radiance, b.radiance), a.key
    print(f"served {st['solves']} solves at {st['solves_per_s']:.1f}/s, one at a time "
          f"{st1['solves_per_s']:.1f}/s")
