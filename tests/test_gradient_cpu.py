"""Pins of the gradient oracle and of the host-side gradient preparation.

* The reference's cost-function known answers (tests/test_derivatives.py:75-135:
  L2 cost 1960.0 and summed derivative -440.0; LL cost 6519.21 and 45329.43),
  reproduced by the oracle's adjoint weights (the adjoint branch contracts
  the same per-pixel terms, so with unit Frechet derivatives the sum of the
  weights is the reference's gradient value).
* The sun-path derivative weights of a uniform layer: the slant optical path
  is linear in the extinction, so d tau / d ext summed over the path is the
  slant length (z_top - z) / mu0, independently of how the path is split.
"""
import numpy as np
import pytest

from at3d_amd import synthetic as S


def test_oracle_cost_function_l2_known_answer(ref):
    stokesout = np.array([10.0, 10.0, 10.0, 0.0])
    w, cost = ref.adjoint_weights(stokesout, np.full(4, 13.0), np.full((4, 4), 5.0), "L2")
    assert cost == pytest.approx(1960.0, abs=1e-5)
    assert w.sum() == pytest.approx(-440.0, abs=1e-5)


def test_oracle_cost_function_ll_known_answer(ref):
    unc = np.zeros((2, 2))
    unc[0, 0] = (1.0 / 0.03) ** 2
    unc[1, 1] = (1.0 / 0.005) ** 2
    w, cost = ref.adjoint_weights(np.array([1.0, 0.5, 0.0]), np.array([1.25, 0.25, 0.25]), unc, "LL")
    assert cost == pytest.approx(6519.21, abs=1e-2)
    assert w.sum() == pytest.approx(45329.43, abs=1e-2)


@pytest.mark.parametrize("mu0", [1.0, 0.5])
def test_oracle_sun_path_weights_sum_to_slant_length(ref, mu0):
    from tests.test_oracle_cpu import _uniform_layer
    w, _ = _uniform_layer(mu0)
    st = ref.setup(ref.scene_from_workload(w))
    ref.init_solution(st)
    shape = w.shape
    ref.prepare_derivatives(st, [ref.Derivative(0, np.ones(shape), np.zeros(shape),
                                                np.zeros((65, 1)), np.ones(shape, np.int32))])
    z = st.gridpos[2, :st.npts]
    np.testing.assert_allclose(st.dpath[:, :st.npts].sum(axis=0), (w.z[-1] - z) / mu0, rtol=1e-5,
                               atol=1e-6)
