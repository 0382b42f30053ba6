"""GPU solve and render against the committed reference outputs in
tests/golden/shdom_small.npz (written by tests/golden/make_golden.py from the
reference Fortran), so parity is checked even where the oracle library is not
built.  Tolerances as for the live comparison: one iteration rel L2 <= 2e-5,
converged fields and rendered radiances rel L2 <= 1e-4, iteration count +-1,
Levis gradient fields rel L2 <= 1e-3 and the cost to 5e-4.
"""
import numpy as np
import pytest

from at3d_amd import gradient as GR
from at3d_amd import synthetic as S
from tests.golden.make_golden import OUT, SCENES, input_checksum
from tests.scenes import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def golden():
    return np.load(OUT)


@pytest.mark.parametrize("name", list(SCENES))
def test_solve_and_render_match_golden(golden, name):
    w = SCENES[name]()
    assert str(golden[f"{name}/checksum"]) == input_checksum(w), "scene generator changed"
    rte = S.make_rte(w)
    rte.solve(1, verbose=False)
    assert rel_l2(rte.state_array("source"), golden[f"{name}/source_iter1"]) < 2e-5
    rte = S.make_rte(w)
    rte.solve(100, verbose=False)
    assert rte.check_solved(verbose=False)
    assert abs(rte.num_iterations - int(golden[f"{name}/iters"])) <= 1
    assert rel_l2(rte.state_array("source"), golden[f"{name}/source"]) < 1e-4
    assert rel_l2(rte.state_array("radiance"), golden[f"{name}/radiance"]) < 1e-4
    assert rel_l2(rte.state_array("fluxes"), golden[f"{name}/fluxes"]) < 1e-4
    for tag in ("views", "interior"):
        rays = {k: golden[f"{name}/{tag}/{k}"] for k in
                ("ray_x", "ray_y", "ray_z", "ray_mu", "ray_phi")}
        got = rte.integrate_to_sensor(rays)["I"]
        assert rel_l2(got, golden[f"{name}/{tag}/I"]) < 1e-4, (tag, rel_l2(got, golden[f"{name}/{tag}/I"]))
    meas = (golden[f"{name}/views/I"] * 1.1).astype(np.float32)
    rays = {k: golden[f"{name}/views/{k}"] for k in ("ray_x", "ray_y", "ray_z", "ray_mu", "ray_phi")}
    sp = list(rte.medium)[0]
    g, c, _ = GR.levis_approximation_grad(rte, dict(rays, measurement_data=meas),
                                          [(sp, "extinction"), (sp, "ssalb")])
    gref = golden[f"{name}/grad"]
    assert c == pytest.approx(float(golden[f"{name}/cost"]), rel=5e-4)
    for i in range(2):
        assert rel_l2(g[i].ravel(), gref[:, i]) < 1e-3, (i, rel_l2(g[i].ravel(), gref[:, i]))
