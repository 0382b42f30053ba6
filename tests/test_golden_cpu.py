"""The committed golden outputs are reproduced by the oracle build and match
the scenes the generators make today."""
import numpy as np
import pytest

from tests.golden.make_golden import OUT, SCENES, dense, input_checksum
from tests.scenes import rel_l2


@pytest.fixture(scope="module")
def golden():
    return np.load(OUT)


@pytest.mark.parametrize("name", list(SCENES))
def test_scene_generators_unchanged(golden, name):
    assert str(golden[f"{name}/checksum"]) == input_checksum(SCENES[name]())


@pytest.mark.parametrize("name", list(SCENES))
def test_oracle_reproduces_golden(ref, golden, name):
    w = SCENES[name]()
    st = ref.setup(ref.scene_from_workload(w))
    ref.init_solution(st)
    ref.solution_iterations(st, 100)
    assert st.iters == int(golden[f"{name}/iters"])
    assert rel_l2(dense(st, "source"), golden[f"{name}/source"]) < 1e-6
    rays = {k: golden[f"{name}/views/{k}"] for k in ("ray_x", "ray_y", "ray_z", "ray_mu", "ray_phi")}
    got = ref.render(st, rays["ray_x"], rays["ray_y"], rays["ray_z"], rays["ray_mu"], rays["ray_phi"])
    assert rel_l2(got, golden[f"{name}/views/I"]) < 1e-6
