"""The per-point medium made on the device (medium.hip) against the host
plan's (the same shared per-point functions, medium_point.h): total
extinction, mixed albedo, source constants, radiance-truncation weights,
single-scatter constants and pairs bit for bit; the two-stream start to
single-precision rounding (it calls exp, whose last bit may differ between
the host's and the device's math libraries); and the solves that follow.
"""
import numpy as np
import pytest

from at3d_amd import synthetic as S
from tests.scenes import blob8, blob8_l15, cloud_l15, cloud_small, rel_l2

pytestmark = pytest.mark.gpu

SCENES = {"blob8_open": lambda: blob8("open"), "blob8_periodic": lambda: blob8("periodic"),
          "cloud12_two_species": cloud_small, "blob8_open_l15": blob8_l15, "cloud16_l15": cloud_l15}


@pytest.mark.parametrize("name", list(SCENES))
def test_device_medium_equals_host_medium(name):
    w = SCENES[name]()
    host = S.make_rte(w, device_setup=False)
    dev = S.make_rte(w, device_setup=True)
    host.solve(0)
    dev.solve(0)
    for k in ("point_ext", "point_alb", "point_chi", "point_tc", "point_tms", "point_ss_off", "point_ss_iph",
              "point_ss_coef", "dirflux", "source", "ns"):
        a, b = host.state_array(k), dev.state_array(k)
        assert a.shape == b.shape and np.array_equal(a, b), k
    r0h, r0d = host.state_array("point_radiance0"), dev.state_array("point_radiance0")
    assert np.allclose(r0d, r0h, rtol=2e-7, atol=0)
    # the plan made on request afterwards holds the same medium
    assert np.array_equal(dev.plan_array("total_ext"), host.plan_array("total_ext"))
    assert np.array_equal(dev.plan_array("chi"), host.plan_array("chi"))
    host.solve(100, verbose=False)
    dev.solve(100, verbose=False)
    assert dev.num_iterations == host.num_iterations
    assert rel_l2(dev.state_array("source"), host.state_array("source")) < 1e-6


def test_device_medium_without_delta_m():
    w = blob8("open")
    w.deltam = False
    host = S.make_rte(w, device_setup=False)
    dev = S.make_rte(w, device_setup=True)
    host.solve(0)
    dev.solve(0)
    for k in ("point_ext", "point_alb", "point_chi", "point_tc", "point_tms", "point_ss_off", "source"):
        assert np.array_equal(host.state_array(k), dev.state_array(k)), k
