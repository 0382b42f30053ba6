"""Several phase tables per property point (num_micro > 1, the reference's
MAXNMICRO): every grid point mixes 8 x num_micro (table, weight) pairs
(TRILIN_INTERP_PROP, shdom90.f90:17-346), which the source, the delta-M
scaling, the two-stream start, the direct beam and the single-scatter
correction all read.  Against the oracle: converged source, radiance and
fluxes, iteration count, and rendered radiances; the device set-up equal to
the host's bit for bit."""
import numpy as np
import pytest

from at3d_amd import synthetic as S
from tests.scenes import micro_blob, ortho_rays, rel_l2

pytestmark = pytest.mark.gpu

SCENES = {"micro2_open": lambda: micro_blob("open"), "micro3_periodic": lambda: micro_blob("periodic", nmicro=3),
          "micro2_open_l15": lambda: micro_blob("open", nmu=16, nphi=32)}


@pytest.mark.parametrize("name", list(SCENES))
def test_micro_solve_and_render_match_reference(ref, name):
    w = SCENES[name]()
    rte = S.make_rte(w)
    rte.solve(100, verbose=False)
    st = ref.setup(ref.scene_from_workload(w))
    ref.init_solution(st)
    ref.solution_iterations(st, 100)
    assert abs(rte.num_iterations - st.iters) <= 1
    n, nlm = st.npts, st.nlm
    assert rel_l2(rte.state_array("source"), ref.dense_sh(st.shptr, st.source, n, nlm)) < 1e-4
    assert rel_l2(rte.state_array("radiance"), ref.dense_sh(st.rshptr, st.radiance, n, nlm)) < 1e-4
    assert rel_l2(rte.state_array("fluxes"), st.fluxes[:, :n]) < 1e-4
    rays = ortho_rays(w, [(0.0, 0.0), (40.0, 120.0)])
    got = rte.integrate_to_sensor(dict(rays))["I"]
    want = ref.render(st, rays["ray_x"], rays["ray_y"], rays["ray_z"], rays["ray_mu"], rays["ray_phi"])
    assert rel_l2(got, want) < 1e-4


@pytest.mark.parametrize("name", list(SCENES)[:2])
def test_micro_device_medium_equals_host(name):
    w = SCENES[name]()
    host = S.make_rte(w, device_setup=False)
    dev = S.make_rte(w, device_setup=True)
    host.solve(0)
    dev.solve(0)
    for k in ("point_ext", "point_alb", "point_chi", "point_tc", "point_tms", "point_ss_off", "point_ss_iph",
              "point_ss_coef", "dirflux", "source"):
        a, b = host.state_array(k), dev.state_array(k)
        assert a.shape == b.shape and np.array_equal(a, b), k
