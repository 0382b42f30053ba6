"""The differentiation workload: the L2 misfit between the 16^3 lognormal
blob's rendered radiances (3 views of 32 x 32 pixels, zenith 0 and +-45
degrees) and a target drawn from the same generator with another seed; the
gradient with respect to per-voxel extinction and single-scattering albedo.

* Against the reference's LEVISAPPROX_GRADIENT on the same inputs: forward
  pixels rel L2 <= 1e-4, cost to 5e-4, each gradient field rel L2 <= 1e-3.
* Against central finite differences (relative step 1e-3) on 32 seeded
  voxels inside the blob: rel L2 <= 5e-2 over the voxels, median per-voxel
  relative difference <= 2e-2 (measured 2.4e-2 / 1.1e-2 for extinction and
  1.0e-2 / 1.3e-3 for albedo), signs equal where the derivative exceeds 5 %
  of the largest.  The Levis approximation differentiates the
  rendered radiance with the spherical-harmonic radiance field held fixed
  (the change of the multiple-scatter field is neglected), so the finite
  differences perturb one voxel, rebuild the source function and direct beam
  of the perturbed medium from the unperturbed converged radiance
  (``RTE.init_solution_from``, no further iteration) and render.  What
  remains between the two is the adjoint's discretisation of the ray
  integral (linear-in-path quadrature of the radiance term, the interpolation
  kernel of the adjoint traversal).
"""
import numpy as np
import pytest

from at3d_amd import gradient as GR
from at3d_amd import synthetic as S
from tests.scenes import rel_l2

pytestmark = [pytest.mark.gpu, pytest.mark.filterwarnings("ignore:RTE is not solved")]

NVOX = 32
STEP = 1e-3


@pytest.fixture(scope="module")
def case():
    w = S.workload("blob16")
    target = S.lognormal_blob(n=16, nleg=64, seed=1)
    rays = S.workload_views("blob16", w)
    t = S.make_rte(target)
    t.solve(target.maxiter, verbose=False)
    meas = t.integrate_to_sensor(dict(rays))["I"].astype(np.float32)
    t.release()
    rte = S.make_rte(w)
    rte.solve(w.maxiter, verbose=False)
    sp = list(rte.medium)[0]
    g, c, out = GR.levis_approximation_grad(rte, dict(rays, measurement_data=meas),
                                            [(sp, "extinction"), (sp, "ssalb")])
    return w, rays, meas, rte, g, c, out["I"]


def test_c3_gradient_matches_reference(ref, case):
    w, rays, meas, rte, g, c, px = case
    st = ref.setup(ref.scene_from_workload(w))
    ref.init_solution(st)
    ref.solution_iterations(st, w.maxiter)
    shape = w.shape
    nleg = w.species[0][2].shape[0]
    ref.prepare_derivatives(st, [
        ref.Derivative(0, np.ones(shape), np.zeros(shape), np.zeros((nleg, 1)), np.ones(shape, np.int32)),
        ref.Derivative(0, np.zeros(shape), np.ones(shape), np.zeros((nleg, 1)), np.ones(shape, np.int32))])
    g_ref, c_ref, px_ref = ref.levis_gradient(st, rays["ray_x"], rays["ray_y"], rays["ray_z"],
                                              rays["ray_mu"], rays["ray_phi"], meas)
    assert rel_l2(px, px_ref) < 1e-4
    assert c == pytest.approx(c_ref, rel=5e-4)
    for i in range(2):
        err = rel_l2(g[i].ravel(), g_ref[:, i])
        print(f"unknown {i}: gradient rel L2 {err:.2e} vs the reference")
        assert err < 1e-3


def _frozen_cost(w, base, rays, meas, ext, alb):
    """L2 cost of the render of (ext, alb) with the source function rebuilt
    from `base`'s converged radiance (no source iteration)."""
    leg = w.species[0][2]
    ww = S.Workload(w.name, w.delx, w.dely, w.z, [(ext.astype(np.float32), alb.astype(np.float32), leg)],
                    bc=w.bc, nmu=w.nmu, nphi=w.nphi, solacc=w.solacc)
    r = S.make_rte(ww)
    r.init_solution_from(base)
    img = r.integrate_to_sensor(dict(rays))["I"].astype(np.float64)
    r.release()
    return 0.5 * np.sum((img - meas) ** 2)


@pytest.mark.parametrize("unknown", [0, 1], ids=["extinction", "ssalb"])
def test_c3_gradient_matches_finite_differences(case, unknown):
    w, rays, meas, rte, g, c, px = case
    ext0 = w.species[0][0].astype(np.float64)
    alb0 = w.species[0][1].astype(np.float64)
    rng = np.random.default_rng(2024)
    inside = np.flatnonzero(ext0.ravel() > 0)
    vox = rng.choice(inside, NVOX, replace=False)
    # the frozen-field model reproduces the converged render at zero step
    c0 = _frozen_cost(w, rte, rays, meas, ext0, alb0)
    assert c0 == pytest.approx(c, rel=1e-4)
    fd = np.zeros(NVOX)
    for k, idx in enumerate(vox):
        i = np.unravel_index(idx, ext0.shape)
        base = ext0 if unknown == 0 else alb0
        h = STEP * base[i]
        cp = []
        for sgn in (1.0, -1.0):
            p = base.copy()
            p[i] += sgn * h
            e, a = (p, alb0) if unknown == 0 else (ext0, p)
            cp.append(_frozen_cost(w, rte, rays, meas, e, a))
        fd[k] = (cp[0] - cp[1]) / (2 * h)
    lev = g[unknown].ravel()[vox]
    err = rel_l2(lev, fd)
    rel = np.abs(lev - fd) / np.maximum(np.abs(fd), 1e-30)
    print(f"{['extinction', 'ssalb'][unknown]}: rel L2 {err:.3e} over {NVOX} voxels, "
          f"median per-voxel {np.median(rel):.2e}, max {rel.max():.2e}")
    for a, b in zip(lev, fd):
        print(f"  levis {a: .6e}  fd {b: .6e}")
    # signs agree wherever the derivative is not negligible (5 % of the
    # largest); a voxel whose derivative is ~1 % of the largest can differ in
    # sign through the adjoint's quadrature of the radiance term
    big = np.abs(fd) > 0.05 * np.abs(fd).max()
    assert np.all(np.sign(lev[big]) == np.sign(fd[big]))
    assert np.median(rel) < 2e-2
    assert err < 5e-2
