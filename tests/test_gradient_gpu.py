"""GPU Levis gradient against the reference (LEVISAPPROX_GRADIENT, adjoint
branch) on the same scenes, rays and measurements.

Both sides start from their own converged solutions (which agree to rel L2
<= 1e-4); forward pixels then agree to 1e-4, the cost to 5e-4 and each
derivative's gradient field to rel L2 <= 1e-3.  A finite-difference check of
the engine's own forward model pins the gradient independently of the
reference in the optically thin limit, where the Levis approximation (source
derivative times transmission, multiple-scatter field changes neglected) is
close to exact: within 20 % at the voxels with the largest gradient (12 % was
observed for the tested blob).
"""
import numpy as np
import pytest

from at3d_amd import gradient as GR
from at3d_amd import synthetic as S
from tests.scenes import blob8, blob8_l15, cloud_l15, cloud_small, ortho_rays, rel_l2

pytestmark = [pytest.mark.gpu, pytest.mark.filterwarnings("ignore:RTE is not solved")]


def _ref_derivs(ref, w, unknowns):
    shape = w.shape
    out = []
    nleg = max(leg.shape[0] for _, _, leg in w.species)
    for isp, var in unknowns:
        ntab = w.species[isp][2].shape[1]
        dext = np.ones(shape) if var == "extinction" else np.zeros(shape)
        dalb = np.ones(shape) if var == "ssalb" else np.zeros(shape)
        out.append(ref.Derivative(isp, dext, dalb, np.zeros((nleg, ntab)), np.ones(shape, np.int32)))
    return out


def _run(ref, w, unknowns, views, cost="L2", ess=True, scale=1.1, splitacc=0.0):
    st = ref.setup(ref.scene_from_workload(w, splitacc=splitacc))
    ref.init_solution(st)
    ref.solution_iterations(st, 100)
    ref.prepare_derivatives(st, _ref_derivs(ref, w, unknowns))
    rays = ortho_rays(w, views)
    base = ref.render(st, rays["ray_x"], rays["ray_y"], rays["ray_z"], rays["ray_mu"], rays["ray_phi"])
    rng = np.random.default_rng(7)
    meas = (base * scale * (1 + 0.05 * rng.standard_normal(base.size))).astype(np.float32)
    meas = np.maximum(meas, 1e-4)
    g_ref, c_ref, px_ref = ref.levis_gradient(st, rays["ray_x"], rays["ray_y"], rays["ray_z"],
                                              rays["ray_mu"], rays["ray_phi"], meas, costfunc=cost,
                                              exact_single_scatter=ess)
    rte = S.make_rte(w, split_accuracy=splitacc)
    rte.solve(100, verbose=False)
    assert rte.state_size("npts") == st.npts
    sensor = dict(rays, measurement_data=meas)
    names = list(rte.medium)
    g, c, out = GR.levis_approximation_grad(rte, sensor, [(names[i], v) for i, v in unknowns],
                                            cost_function=cost, exact_single_scatter=ess)
    return g, c, out["I"], g_ref, c_ref, px_ref


def _check(res, tol=1e-3):
    g, c, px, g_ref, c_ref, px_ref = res
    assert rel_l2(px, px_ref) < 1e-4
    assert c == pytest.approx(c_ref, rel=5e-4)
    for i in range(g.shape[0]):
        gi = g[i].ravel()
        assert np.all(np.isfinite(gi))
        assert rel_l2(gi, g_ref[:, i]) < tol, (i, rel_l2(gi, g_ref[:, i]))


VIEWS = [(0.0, 0.0), (45.0, 30.0), (60.0, 200.0)]


@pytest.mark.parametrize("bc", ["open", "periodic"])
def test_extinction_and_albedo_gradient_matches_reference(ref, bc):
    _check(_run(ref, blob8(bc), [(0, "extinction"), (0, "ssalb")], VIEWS))


@pytest.mark.parametrize("bc", ["open", "periodic"])
def test_gradient_on_refined_grid_matches_reference(ref, bc):
    """Adaptive cell splitting (split_accuracy 0.03) grows the grid 4-5x; the
    adjoint rays then cross split cells (tree descents, faces onto smaller
    cells) and the sun paths start from the new points."""
    _check(_run(ref, blob8(bc), [(0, "extinction"), (0, "ssalb")], VIEWS, splitacc=0.03))


def test_two_species_gradient_matches_reference(ref):
    _check(_run(ref, cloud_small(), [(0, "extinction"), (1, "extinction"), (0, "ssalb")],
                VIEWS[:2]))


def test_two_species_l15_gradient_matches_reference(ref):
    """NMU 16 / NPHI 32 (L = 15): the 200-term cloud table (delta-M, full-table
    single-scatter correction) and Rayleigh; extinction of both and the cloud
    albedo."""
    _check(_run(ref, cloud_l15(), [(0, "extinction"), (1, "extinction"), (0, "ssalb")], VIEWS[:2]))


def test_blob_l15_gradient_matches_reference(ref):
    _check(_run(ref, blob8_l15("periodic"), [(0, "extinction"), (0, "ssalb")], VIEWS))


def test_gradient_without_exact_single_scatter_matches_reference(ref):
    _check(_run(ref, blob8("open"), [(0, "extinction")], VIEWS[:2], ess=False))


def test_log_cost_gradient_matches_reference(ref):
    _check(_run(ref, blob8("open"), [(0, "extinction"), (0, "ssalb")], VIEWS[:2], cost="LL"))


def test_phase_legendre_gradient_matches_reference(ref):
    """d/d chi_l of the phase table (the 'legendre_0_l' unknowns)."""
    w = blob8("open")
    unknowns = [(0, "legendre_0_3")]
    st = ref.setup(ref.scene_from_workload(w))
    ref.init_solution(st)
    ref.solution_iterations(st, 100)
    nleg = w.species[0][2].shape[0]
    dleg = np.zeros((nleg, 1))
    dleg[3, 0] = 1.0
    shape = w.shape
    ref.prepare_derivatives(st, [ref.Derivative(0, np.zeros(shape), np.zeros(shape), dleg,
                                                np.ones(shape, np.int32))])
    rays = ortho_rays(w, VIEWS[:2])
    base = ref.render(st, rays["ray_x"], rays["ray_y"], rays["ray_z"], rays["ray_mu"], rays["ray_phi"])
    meas = (base * 0.9).astype(np.float32)
    g_ref, c_ref, px_ref = ref.levis_gradient(st, rays["ray_x"], rays["ray_y"], rays["ray_z"],
                                              rays["ray_mu"], rays["ray_phi"], meas)
    rte = S.make_rte(w)
    rte.solve(100, verbose=False)
    g, c, out = GR.levis_approximation_grad(rte, dict(rays, measurement_data=meas),
                                            [(list(rte.medium)[0], "legendre_0_3")])
    _check((g, c, out["I"], g_ref, c_ref, px_ref))


def test_gradient_matches_finite_differences_in_thin_limit():
    """Thin, single-scattering-dominated blob: the Levis gradient of the L2
    cost with respect to extinction agrees with central differences of the
    engine's own forward model (solve + render) at the voxels that matter."""
    w = blob8("open")
    ext0 = w.species[0][0] * 0.02          # column optical depth ~ 0.03
    alb0 = w.species[0][1]
    leg = w.species[0][2]
    rays = ortho_rays(w, [(0.0, 0.0), (45.0, 30.0)])

    def forward(ext):
        ww = S.Workload(w.name, w.delx, w.dely, w.z, [(ext.astype(np.float32), alb0, leg)], bc=w.bc,
                        nmu=w.nmu, nphi=w.nphi, solacc=1e-6)
        rte = S.make_rte(ww)
        rte.solve(100, verbose=False)
        return rte, rte.integrate_to_sensor(dict(rays))["I"].astype(np.float64)

    rte, base = forward(ext0)
    meas = (base * 0.8).astype(np.float32)
    g, c, _ = GR.levis_approximation_grad(rte, dict(rays, measurement_data=meas),
                                          [(list(rte.medium)[0], "extinction")])
    g = g[0]
    # voxels inside the medium with the largest gradient magnitude
    cand = np.where(ext0.ravel() > 0.05 * ext0.max(), np.abs(g.ravel()), -1.0)
    flat = np.argsort(-cand)[:6]
    h = 0.05
    for idx in flat:
        i = np.unravel_index(idx, g.shape)
        de = ext0[i] * h
        ep, em = ext0.copy(), ext0.copy()
        ep[i] += de
        em[i] -= de
        cp = 0.5 * np.sum((forward(ep)[1] - meas) ** 2)
        cm = 0.5 * np.sum((forward(em)[1] - meas) ** 2)
        fd = (cp - cm) / (2 * de)
        assert g[i] == pytest.approx(fd, rel=0.2, abs=1e-3 * np.max(np.abs(g))), (i, g[i], fd)


def test_levis_loss_autograd():
    """torch.autograd.Function: backward returns the Levis gradient normalised
    by the number of measurements, forward the normalised cost."""
    import torch
    w = blob8("open")
    rays = ortho_rays(w, [(0.0, 0.0)])
    rte0 = S.make_rte(w)
    rte0.solve(100, verbose=False)
    meas = (rte0.integrate_to_sensor(dict(rays))["I"] * 1.2).astype(np.float32)
    sensor = dict(rays, measurement_data=meas)
    leg = w.species[0][2]

    def make(ext, alb):
        ww = S.Workload(w.name, w.delx, w.dely, w.z, [(ext.astype(np.float32), alb.astype(np.float32), leg)],
                        bc=w.bc, nmu=w.nmu, nphi=w.nphi, solacc=w.solacc)
        return S.make_rte(ww)

    ext = torch.tensor(w.species[0][0], dtype=torch.float64, requires_grad=True)
    alb = torch.tensor(w.species[0][1], dtype=torch.float64, requires_grad=True)
    loss = GR.levis_loss(ext, alb, make, [sensor], list(rte0.medium)[0])
    loss.backward()
    g, c, _ = GR.levis_approximation_grad(rte0, sensor, [(list(rte0.medium)[0], "extinction"),
                                                         (list(rte0.medium)[0], "ssalb")])
    n = meas.size
    assert float(loss) == pytest.approx(c / n, rel=1e-5)
    assert rel_l2(ext.grad.numpy(), g[0] / n) < 1e-5
    assert rel_l2(alb.grad.numpy(), g[1] / n) < 1e-5


def _jacobian_case(ref, w, unknowns, views, points, rays_per_pixel=1, seed=11):
    st = ref.setup(ref.scene_from_workload(w))
    ref.init_solution(st)
    ref.solution_iterations(st, 100)
    ref.prepare_derivatives(st, _ref_derivs(ref, w, unknowns))
    rays = ortho_rays(w, views)
    n = rays["ray_x"].size
    if rays_per_pixel > 1:
        # each pixel averages rays_per_pixel rays shifted by a fraction of a cell
        rng = np.random.default_rng(seed)
        k = rays_per_pixel
        rays = {key: np.repeat(v, k) for key, v in rays.items()}
        rays["ray_x"] = (rays["ray_x"] + rng.uniform(-0.3, 0.3, n * k) * w.delx).astype(np.float32)
        rays["ray_y"] = (rays["ray_y"] + rng.uniform(-0.3, 0.3, n * k) * w.dely).astype(np.float32)
    rpp = np.full(n, rays_per_pixel, np.int32)
    rw = np.full(n * rays_per_pixel, 1.0 / rays_per_pixel)
    args = (rays["ray_x"], rays["ray_y"], rays["ray_z"], rays["ray_mu"], rays["ray_phi"])
    img = ref.render(st, *args)
    meas = (img.reshape(n, rays_per_pixel).mean(axis=1) * 1.1).astype(np.float32)
    flat = np.ravel_multi_index(points, w.shape)
    g_ref, c_ref, px_ref, J_ref = ref.levis_gradient(st, *args, meas, rays_per_pixel=rpp, ray_weights=rw,
                                                     jacobian_ptr=flat)
    rte = S.make_rte(w)
    rte.solve(100, verbose=False)
    sensor = dict(rays, measurement_data=meas, rays_per_pixel=rpp, ray_weight=rw)
    names = list(rte.medium)
    g, c, out, jac = GR.levis_approximation_grad(rte, sensor, [(names[i], v) for i, v in unknowns],
                                                 indices_for_jacobian=points)
    return g, c, out["I"], jac, g_ref, c_ref, px_ref, J_ref


@pytest.mark.parametrize("bc", ["open", "periodic"])
def test_jacobian_matches_reference(ref, bc):
    """Jacobian branch (MAKEJACOBIAN): per-pixel derivatives at chosen property
    points inside, at the edge of and outside the blob, for extinction and
    albedo, against the reference's JACOBIAN; rel L2 <= 1e-3 per unknown."""
    w = blob8(bc)
    points = (np.array([3, 4, 0, 5, 7, 4]), np.array([4, 3, 0, 5, 7, 4]), np.array([4, 2, 0, 3, 7, 5]))
    g, c, px, jac, g_ref, c_ref, px_ref, J_ref = _jacobian_case(
        ref, w, [(0, "extinction"), (0, "ssalb")], VIEWS, points)
    assert jac.shape == (1, 2, 6, px.size)
    J = np.transpose(jac[0], (2, 1, 0))                  # (npix, njac, numder)
    assert rel_l2(px, px_ref) < 1e-4
    assert c == pytest.approx(c_ref, rel=5e-4)
    for d in range(2):
        assert np.abs(J_ref[:, :, d]).max() > 0
        assert rel_l2(J[:, :, d], J_ref[:, :, d]) < 1e-3, (d, rel_l2(J[:, :, d], J_ref[:, :, d]))
        assert rel_l2(g[d].ravel(), g_ref[:, d]) < 1e-3


def test_jacobian_with_several_rays_per_pixel_matches_reference(ref):
    """Two species, two rays per pixel averaged with weights 1/2: the
    Jacobian sums the weighted rays of each pixel as RAYGRAD_PIXEL does."""
    w = cloud_small()
    pts = (np.array([6, 5, 0, 7]), np.array([6, 6, 3, 2]), np.array([3, 4, 9, 5]))
    g, c, px, jac, g_ref, c_ref, px_ref, J_ref = _jacobian_case(
        ref, w, [(0, "extinction"), (1, "extinction")], VIEWS[:2], pts, rays_per_pixel=2)
    J = np.transpose(jac[0], (2, 1, 0))
    assert rel_l2(px, px_ref) < 1e-4
    for d in range(2):
        assert rel_l2(J[:, :, d], J_ref[:, :, d]) < 1e-3, (d, rel_l2(J[:, :, d], J_ref[:, :, d]))


def test_levis_approx_gradient_uncorrelated_over_two_wavelengths(ref):
    """LevisApproxGradientUncorrelated (gradient.py:417-449) with two solvers
    keyed by wavelength and one instrument holding a sensor of each: the
    loss and gradient are the per-sensor reference gradients summed and
    divided by the number of measurements; the Jacobian dataset holds each
    wavelength's pixels."""
    base = blob8("open")
    ws = {}
    for wl, view in ((0.86, (0.0, 0.0)), (0.66, (45.0, 30.0))):
        ext = (base.species[0][0] * (0.86 / wl)).astype(np.float32)
        ws[wl] = (S.Workload(base.name, base.delx, base.dely, base.z,
                             [(ext, base.species[0][1], base.species[0][2])], bc=base.bc,
                             nmu=base.nmu, nphi=base.nphi, solacc=base.solacc, wavelength=wl), view)
    solvers, meas_list, fwd_list = {}, [], []
    g_sum, c_sum, n = None, 0.0, 0
    for wl, (w, view) in ws.items():
        st = ref.setup(ref.scene_from_workload(w))
        ref.init_solution(st)
        ref.solution_iterations(st, 100)
        ref.prepare_derivatives(st, _ref_derivs(ref, w, [(0, "extinction"), (0, "ssalb")]))
        rays = ortho_rays(w, [view])
        args = [rays[k] for k in ("ray_x", "ray_y", "ray_z", "ray_mu", "ray_phi")]
        m = (ref.render(st, *args) * 1.15).astype(np.float32)
        g_ref, c_ref, _ = ref.levis_gradient(st, *args, m)
        g_sum = g_ref if g_sum is None else g_sum + g_ref
        c_sum += c_ref
        n += m.size
        rte = S.make_rte(w)
        rte.solve(100, verbose=False)
        solvers[wl] = rte
        meas_list.append(dict(rays, measurement_data=m, wavelength=wl))
        fwd_list.append(dict(rays, wavelength=wl))
    name = list(solvers[0.86].medium)[0]
    jac_pts = (np.array([3, 4]), np.array([4, 4]), np.array([3, 5]))
    lg = GR.LevisApproxGradientUncorrelated(
        {"cam": {"sensor_list": meas_list, "uncertainty_model": None}}, solvers,
        {"cam": {"sensor_list": fwd_list}}, {name: ["extinction", "ssalb"]},
        {"maxiter": 100, "verbose": False, "n_jobs": 4},
        {"cost_function": "L2", "exact_single_scatter": True, "indices_for_jacobian": jac_pts},
        {"add_noise": False})
    loss, gds, jds = lg()
    assert loss == pytest.approx(c_sum / n, rel=5e-4)
    grad = gds["gradient"]
    assert grad.shape == base.shape + (2,)
    assert gds["derivative_index"] == [(name, "extinction"), (name, "ssalb")]
    for i in range(2):
        err = rel_l2(grad[..., i].ravel(), g_sum[:, i] / n)
        print(f"derivative {i}: rel L2 {err:.2e} vs the summed reference gradients")
        assert err < 1e-3
    for k, (wl, (w, view)) in enumerate(ws.items()):
        assert np.all(np.isfinite(fwd_list[k]["I"])) and fwd_list[k]["I"].size == meas_list[k]["ray_x"].size
        jac = jds["jacobian_{:1.3f}".format(wl)]
        assert jac.shape == (1, 2, 2, meas_list[k]["ray_x"].size)
        assert np.any(jac != 0)
    assert len(jds["grid_index"]) == 2
