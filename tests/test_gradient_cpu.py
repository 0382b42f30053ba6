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


def test_oracle_jacobian_branch_is_the_per_pixel_derivative(ref):
    """The oracle's Jacobian branch (GRAD_INTEGRATE_1RAY + UPDATE_COSTFUNCTION)
    against its adjoint branch: the two gradients agree (rel 1e-5), and at the
    Jacobian's points the L2 gradient is sum_pixels (R - M) x Jacobian (the
    chain rule that defines the Jacobian), exactly up to rounding."""
    from tests.scenes import blob8, ortho_rays, rel_l2
    w = blob8("open")
    st = ref.setup(ref.scene_from_workload(w))
    ref.init_solution(st)
    ref.solution_iterations(st, 100)
    shape = w.shape
    ref.prepare_derivatives(st, [ref.Derivative(0, np.ones(shape), np.zeros(shape), np.zeros((65, 1)),
                                                np.ones(shape, np.int32)),
                                 ref.Derivative(0, np.zeros(shape), np.ones(shape), np.zeros((65, 1)),
                                                np.ones(shape, np.int32))])
    rays = ortho_rays(w, [(0.0, 0.0), (45.0, 30.0)])
    args = (rays["ray_x"], rays["ray_y"], rays["ray_z"], rays["ray_mu"], rays["ray_phi"])
    img = ref.render(st, *args)
    meas = (img * 1.1).astype(np.float32)
    g, c, px = ref.levis_gradient(st, *args, meas)
    jp = np.array([0, 100, 219, 300, 511])
    gj, cj, pxj, J = ref.levis_gradient(st, *args, meas, jacobian_ptr=jp)
    assert J.shape == (px.size, jp.size, 2)
    assert cj == pytest.approx(c, rel=1e-5)
    for i in range(2):
        assert rel_l2(gj[:, i], g[:, i]) < 1e-5
    gsel = np.einsum("p,pjd->jd", (pxj - meas).astype(np.float64), J.astype(np.float64))
    np.testing.assert_allclose(gsel, gj[jp], rtol=1e-5, atol=1e-9 * np.abs(gj).max())
    assert np.abs(J[:, 2, 1]).max() > 0   # a point inside the blob responds to ssalb


def test_levis_approx_gradient_argument_checks():
    """The reference's constructor checks (gradient.py:61-127): a missing
    cost function, a non-boolean 'add_noise', an uncertainty model for
    another cost function; NullUncertainty fills unit inverse covariances."""
    import warnings

    from at3d_amd import gradient as GR
    sens = {"measurement_data": np.ones(5, np.float32), "wavelength": 0.86}
    meas = lambda model=None: {"cam": {"sensor_list": [dict(sens)], "uncertainty_model": model}}
    with pytest.raises(ValueError, match="cost_function"):
        GR.LevisApproxGradientUncorrelated(meas(), {}, {}, {}, {}, {}, {"add_noise": False})
    with pytest.raises(TypeError, match="add_noise"):
        GR.LevisApproxGradientUncorrelated(meas(), {}, {}, {}, {}, {"cost_function": "L2"},
                                           {"add_noise": 1})
    with pytest.raises(ValueError, match="inconsistent"):
        GR.LevisApproxGradientUncorrelated(meas(GR.NullUncertainty("LL")), {}, {}, {}, {},
                                           {"cost_function": "L2"}, {"add_noise": False})
    m = meas()
    with warnings.catch_warnings(record=True) as rec:
        warnings.simplefilter("always")
        lg = GR.LevisApproxGradientUncorrelated(m, {}, {}, {}, {}, {"cost_function": "L2"}, {})
    assert any("add_noise" in str(r.message) for r in rec)
    assert any("NullUncertainty" in str(r.message) for r in rec)
    assert np.array_equal(m["cam"]["sensor_list"][0]["uncertainties"], np.ones(5))
    assert lg.nmeasurements == 5
