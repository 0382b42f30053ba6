"""The polarized (NSTOKES = 3) solve and render on the GPU against the
reference's polarized SHDOM (the oracle, same inputs): the source function's
I, Q and U terms after INIT_SOLUTION, after one iteration and at convergence
(with the iteration count), the hemispheric fluxes, and the rendered I, Q, U
of ortho, interior and up-looking rays; plus the a1-only equivalence with the
scalar GPU solve.

Tolerances as the scalar tests: the reference sums its dot products in single
precision and transforms with FFT/DFT, this engine with FP64 reductions and
FP32 MFMA GEMMs; one iteration agrees to rel L2 <= 2e-5, converged fields to
1e-4.  Rendered Q and U are checked to 1e-4 of the largest |I| plus 1e-3 of
themselves (they are one to two orders below I).
"""
import numpy as np
import pytest

from at3d_amd import synthetic as S
from tests.scenes import interior_rays, ortho_rays, pol_blob, rel_l2

pytestmark = pytest.mark.gpu

SCENES = {"pol_blob8_open": lambda: pol_blob(), "pol_blob8_periodic": lambda: pol_blob(bc="periodic"),
          "pol_blob8_open_l15": lambda: pol_blob(nmu=16, nphi=32)}


def _ref_solve(ref, w, maxiter):
    st = ref.setup(ref.scene_from_workload(w))
    ref.init_solution(st)
    if maxiter > 0:
        ref.solution_iterations(st, maxiter)
    return st


def _ref_fields(ref, st, which):
    ptr, arr = (st.shptr, st.source) if which == "source" else (st.rshptr, st.radiance)
    return np.stack([ref.dense_sh(ptr, arr, st.npts, st.nlm, k) for k in range(3)])


@pytest.mark.parametrize("name", list(SCENES))
def test_polarized_source_init_and_one_iteration(ref, name):
    w = SCENES[name]()
    rte = S.make_rte(w)
    rte.solve(0)
    st = _ref_solve(ref, w, 0)
    assert np.array_equal(rte.state_array("ns"), np.diff(st.shptr[:st.npts + 1]))
    got, want = rte.state_array("source"), _ref_fields(ref, st, "source")
    for k in range(3):
        assert rel_l2(got[k], want[k]) < 1e-6, k
    assert np.abs(want[1]).max() > 0.0          # the first source is polarized
    rte.solve(1, init_solution=False)
    ref.solution_iterations(st, 1)
    got, want = rte.state_array("source"), _ref_fields(ref, st, "source")
    for k in range(3):
        assert rel_l2(got[k], want[k]) < 2e-5, k
    assert np.abs(want[2]).max() > 0.0          # U appears after one scattering


@pytest.mark.parametrize("name", list(SCENES))
def test_polarized_converged_fields(ref, name):
    w = SCENES[name]()
    rte = S.make_rte(w)
    rte.solve(100, verbose=False)
    st = _ref_solve(ref, w, 100)
    assert abs(rte.num_iterations - st.iters) <= 1
    for which in ("source", "radiance"):
        got, want = rte.state_array(which), _ref_fields(ref, st, which)
        for k in range(3):
            assert rel_l2(got[k], want[k]) < 1e-4, (which, k)
    n = st.npts
    assert rel_l2(rte.state_array("fluxes"), st.fluxes[:, :n]) < 1e-4


@pytest.mark.parametrize("name", list(SCENES))
def test_polarized_render(ref, name):
    w = SCENES[name]()
    rte = S.make_rte(w)
    rte.solve(100, verbose=False)
    st = _ref_solve(ref, w, 100)
    rays = ortho_rays(w, [(0.0, 0.0), (35.0, 60.0), (60.0, 200.0)])
    inner = interior_rays(w, 64, seed=4)
    for key in rays:
        rays[key] = np.concatenate([rays[key], inner[key]])
    out = rte.integrate_to_sensor(dict(rays))
    want = ref.render(st, rays["ray_x"], rays["ray_y"], rays["ray_z"], rays["ray_mu"], rays["ray_phi"])
    imax = np.abs(want[0]).max()
    assert rel_l2(out["I"], want[0]) < 1e-4
    for k, c in ((1, "Q"), (2, "U")):
        assert np.allclose(out[c], want[k], rtol=1e-3, atol=1e-4 * imax), c
        assert np.abs(want[k]).max() > 1e-3 * imax, c     # the check has something to see


def test_polarized_single_scatter_render(ref):
    w = pol_blob()
    rte = S.make_rte(w)
    rte.solve(100, verbose=False)
    st = _ref_solve(ref, w, 100)
    rays = ortho_rays(w, [(20.0, 30.0)])
    out = rte.integrate_to_sensor(dict(rays), single_scatter=True)
    want = ref.render(st, rays["ray_x"], rays["ray_y"], rays["ray_z"], rays["ray_mu"], rays["ray_phi"],
                      single_scatter=True)
    imax = np.abs(want[0]).max()
    assert rel_l2(out["I"], want[0]) < 1e-4
    for k, c in ((1, "Q"), (2, "U")):
        assert np.allclose(out[c], want[k], rtol=1e-3, atol=1e-4 * imax), c


def test_a1_only_polarized_equals_scalar():
    """No polarizing element: Q and U stay zero and I is the scalar solve's."""
    w = pol_blob()
    species = []
    for e, a, leg in w.species:
        t = np.zeros_like(leg)
        t[0] = leg[0]
        species.append((e, a, t))
    w3 = S.Workload("a1_only", w.delx, w.dely, w.z, species, bc=w.bc, nmu=w.nmu, nphi=w.nphi, solacc=w.solacc,
                    nstokes=3)
    w1 = S.Workload("a1_only", w.delx, w.dely, w.z, [(e, a, t[0]) for e, a, t in species], bc=w.bc, nmu=w.nmu,
                    nphi=w.nphi, solacc=w.solacc)
    r3, r1 = S.make_rte(w3), S.make_rte(w1)
    r3.solve(100, verbose=False)
    r1.solve(100, verbose=False)
    assert abs(r3.num_iterations - r1.num_iterations) <= 1
    s3 = r3.state_array("source")
    assert np.all(s3[1:] == 0.0)
    rays = ortho_rays(w, [(0.0, 0.0), (45.0, 120.0)])
    o3 = r3.integrate_to_sensor(dict(rays))
    o1 = r1.integrate_to_sensor(dict(rays))
    assert rel_l2(o3["I"], o1["I"]) < 1e-5
    assert np.all(o3["Q"] == 0.0) and np.all(o3["U"] == 0.0)


@pytest.mark.parametrize("bc", ["open", "periodic"])
def test_polarized_point_ordered_sweep(ref, bc):
    """Grids whose rings do not fit in LDS sweep point by point: the same I,
    Q, U (the one-cell path bit for bit against the general trace) and the
    reference's converged fields."""
    w = pol_blob(bc=bc)
    a = S.make_rte(w)
    a.set_option("sweep_ring", 0)
    a.solve(3, verbose=False)
    b = S.make_rte(w)
    b.set_option("sweep_general", 1)
    b.solve(3, verbose=False)
    assert a.timing("sweep_ring") == 0.0
    assert np.array_equal(a.state_array("source"), b.state_array("source"))
    c = S.make_rte(w)
    c.set_option("sweep_ring", 0)
    c.solve(100, verbose=False)
    st = _ref_solve(ref, w, 100)
    assert abs(c.num_iterations - st.iters) <= 1
    got, want = c.state_array("source"), _ref_fields(ref, st, "source")
    for k in range(3):
        assert rel_l2(got[k], want[k]) < 1e-4, k


@pytest.mark.parametrize("bc", ["open", "periodic"])
def test_polarized_ring_of_radiances(bc):
    """The 4-byte ring (radiances alone) against the ring of pairs for I, Q
    and U after three iterations (same cell integrals from the same values)."""
    w = pol_blob(bc=bc)
    a = S.make_rte(w)
    a.set_option("sweep_ring", 2)
    a.solve(3, verbose=False)
    b = S.make_rte(w)
    b.set_option("sweep_ring", 2)
    b.set_option("sweep_ring4", 1)
    b.solve(3, verbose=False)
    assert a.timing("sweep_ring") == 1.0 and b.timing("sweep_ring4") == 1.0
    for name in ("source", "radiance"):
        got, want = b.state_array(name), a.state_array(name)
        for k in range(3):
            assert rel_l2(got[k], want[k]) < 1e-7, (name, k)


def test_polarized_large_grid_takes_ring_of_radiances():
    """A 96 x 96 x 64 polarized grid (16 x 32 ordinates), whose ring of
    (radiance, source) pairs needs about 200 KB of LDS, takes the 4-byte ring
    (about 100 KB) by itself; two iterations of it agree with the
    point-ordered sweep (rel L2 <= 2e-6 for I, Q, U)."""
    w = S.polarized_cloud_field(nx=96, ny=96, nz=64, seed=2)
    a = S.make_rte(w, max_total_mb=60000, adapt_grid_factor=1.0)
    a.solve(2, verbose=False)
    assert a.timing("sweep_ring") == 1.0 and a.timing("sweep_ring4") == 1.0
    b = S.make_rte(w, max_total_mb=60000, adapt_grid_factor=1.0)
    b.set_option("sweep_ring", 0)
    b.solve(2, verbose=False)
    got, want = a.state_array("source"), b.state_array("source")
    for k in range(3):
        assert rel_l2(got[k], want[k]) < 2e-6, k
