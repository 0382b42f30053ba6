"""SolversDict / SensorsDict (at3d/containers.py): two wavelengths solved on
two host threads give the same solutions as solved one by one, and
get_measurements stores each sensor's pixels as the solver of its wavelength
renders them (sub-pixel rays averaged with their weights)."""
import dataclasses

import numpy as np
import pytest

from at3d_amd import sensor as SE
from at3d_amd import synthetic as S
from at3d_amd.containers import SensorsDict, SolversDict
from tests.scenes import blob8

pytestmark = [pytest.mark.gpu, pytest.mark.filterwarnings("ignore:RTE is not solved")]


def _solvers(wls):
    base = blob8("open")
    sd = SolversDict()
    for wl in wls:
        w = dataclasses.replace(base, wavelength=wl)
        sd.add_solver(wl, S.make_rte(w))
    return base, sd


def test_threaded_solves_and_measurements():
    wls = (0.66, 0.86)
    base, sd = _solvers(wls)
    sd.solve(n_jobs=2, verbose=False)
    assert all(s.check_solved(verbose=False) for s in sd.values())
    assert sd.to_solve(False)[0] == []
    _, one = _solvers(wls)
    one.solve(n_jobs=1, verbose=False)
    for wl in wls:
        assert np.array_equal(sd[wl].state_array("source"), one[wl].state_array("source"))
    sensors = SensorsDict()
    box = (0.0, 0.0, 0.0, 0.35, 0.35, 0.28)
    for wl in wls:
        cam = SE.orthographic_projection(wl, box, 0.05, 0.05, 0.0, 0.0, altitude="TOA", stokes=["I"],
                                         sub_pixel_ray_args={"method": SE.gaussian, "degree": (2, 2)})
        sensors.add_sensor("cam", cam)
    assert sensors.get_unique_solvers() == sorted(wls)
    sensors.get_measurements(sd, n_jobs=2, verbose=False)
    for s in sensors["cam"]["sensor_list"]:
        rays = {k: s[k] for k in ("ray_x", "ray_y", "ray_z", "ray_mu", "ray_phi")}
        want = SE.average_subpixel_rays(s, one[s["wavelength"]].integrate_to_sensor(rays)["I"])
        assert s["I"].shape == (np.size(s["cam_x"]),)
        assert np.allclose(s["I"], want, rtol=1e-6, atol=0)
    with pytest.raises(TypeError):
        sd.add_solver(1.0, object())
