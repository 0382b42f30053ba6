"""Pins of the oracle (the reference Fortran driven through ctypes) and of the
host setup against what the mathematics fixes, independently of either code:

* Gauss-Legendre quadrature (numpy.polynomial.legendre.leggauss) for the
  discrete-ordinate cosines and weights, and the ordinate counts 90 / 354 that
  the reference's own original-SHDOM output files print for NMU 8 / NPHI 16
  and NMU 16 / NPHI 32 (tests/data/brdf_L1r.out:2 and rico32x36x26w672ar.out:2);
* the direct beam of a horizontally uniform absorbing layer, F0 exp(-tau/mu0);
* the closed form of a pure absorber over a Lambertian surface seen at nadir
  with an overhead sun, I = exp(-tau) * albedo / pi * exp(-tau)
  (the reference's test_shdom.py:855-908 checks its solver the same way);
* SH -> DO -> SH is the identity where the quadrature is exact (full Gaussian
  angle set; the m = 0 terms of the reduced set).
A mistake such as a dropped term, a wrong sign or a transposed table fails one
of these.
"""
import numpy as np
import pytest

from at3d_amd import synthetic as S
from tests.scenes import absorber_columns, rel_l2


def _nadir_rays(w):
    nx, ny, _ = w.shape
    x = np.arange(nx) * w.delx
    y = np.full(nx, (ny // 2) * w.dely)
    z = np.full(nx, float(w.z[-1]))
    return x, y, z, np.ones(nx), np.zeros(nx)


@pytest.mark.parametrize("nmu", [8, 16])
def test_oracle_quadrature_is_gauss_legendre(ref, nmu):
    """NMU-point Gauss-Legendre on [-1, 1]: down-welling cosines first, in
    increasing order, then the up-welling ones mirrored; weights sum to 2 and
    the discrete-ordinate weights to 4 pi."""
    a = ref.make_angle_set(nmu, 2 * nmu, 2, 2 * nmu)
    t, wt = np.polynomial.legendre.leggauss(nmu)
    h = nmu // 2
    np.testing.assert_allclose(a["mu"][:h], t[:h], rtol=1e-6)
    np.testing.assert_allclose(a["mu"][h:], -t[:h], rtol=1e-6)
    np.testing.assert_allclose(a["wtmu"][:h], wt[:h], rtol=3e-6)
    np.testing.assert_allclose(a["wtmu"][h:], wt[:h], rtol=3e-6)
    assert abs(float(np.sum(a["wtdo"], dtype=np.float64)) - 4 * np.pi) < 1e-5


@pytest.mark.parametrize("nmu,nphi,nang", [(8, 16, 90), (16, 32, 354)])
def test_oracle_and_plan_ordinate_counts(ref, nmu, nphi, nang):
    assert ref.make_angle_set(nmu, nphi, 2, nphi)["nang"] == nang
    w = S.lognormal_blob(n=8, nmu=nmu, nphi=nphi, nleg=64)
    assert S.make_rte(w).plan_size("nang") == nang


def _uniform_layer(mu0, ext=1.5):
    nx = ny = 4
    z = np.linspace(0.0, 1.0, 6).astype(np.float32)
    e = np.full((nx, ny, z.size), ext, np.float32)
    a = np.zeros_like(e)
    return S.Workload("uniform_layer", 0.1, 0.1, z, [(e, a, S.hg_legendre(0.0, 0))], bc="periodic",
                      nmu=8, nphi=16, solarmu=-mu0, solaraz_deg=30.0), ext


@pytest.mark.parametrize("mu0", [1.0, 0.5])
def test_oracle_direct_beam_uniform_layer(ref, mu0):
    w, ext = _uniform_layer(mu0)
    st = ref.setup(ref.scene_from_workload(w))
    ref.init_solution(st)
    npts = st.npts
    z = st.gridpos[2, :npts]
    want = np.exp(-ext * (float(w.z[-1]) - z) / mu0)
    np.testing.assert_allclose(st.dirflux[:npts], want, rtol=2e-6)


@pytest.mark.parametrize("mu0", [1.0, 0.5])
def test_plan_direct_beam_uniform_layer(mu0):
    w, ext = _uniform_layer(mu0)
    rte = S.make_rte(w)
    z = rte.plan_array("gridpos")[:, 2]
    want = np.exp(-ext * (float(w.z[-1]) - z) / mu0)
    np.testing.assert_allclose(rte.plan_array("dirflux"), want, rtol=2e-6)


@pytest.mark.parametrize("bc", ["open", "periodic"])
def test_oracle_absorber_closed_form(ref, bc):
    w, tau = absorber_columns(bc=bc)
    st = ref.setup(ref.scene_from_workload(w))
    ref.init_solution(st)
    ref.solution_iterations(st, 10)
    got = ref.render(st, *_nadir_rays(w))
    want = np.exp(-tau) * 0.04 / np.pi * np.exp(-tau)
    # the render sub-step transmission is a 3-term Taylor series of exp(-dtau)
    # with dtau <= 0.2, a relative error below 1e-4 over tau <= 2
    np.testing.assert_allclose(got, want, rtol=3e-4)


def test_plan_sh_do_round_trip_is_identity_with_full_gauss_angles():
    """With the full Gaussian angle set (NPHI0 = NPHI at every cosine) the
    quadrature integrates every product Y_J Y_J' of the basis exactly (degree
    <= 2 ML = 2 NMU - 2 in mu, |m| + |m'| < NPHI in azimuth), so analysing a
    synthesised field returns its coefficients: W^T Y = I."""
    w = S.lognormal_blob(n=8, nmu=8, nphi=16, nleg=64)
    rte = S.make_rte(w, angle_set=1)
    y = rte.plan_array("ymat").astype(np.float64)   # [nang][nlm]
    wm = rte.plan_array("wmat").astype(np.float64)
    assert rte.plan_size("nang") == 8 * 16
    g = wm.T @ y
    assert np.max(np.abs(g - np.eye(g.shape[0]))) < 2e-6


def test_plan_reduced_gauss_round_trip_exact_on_low_azimuth_terms():
    """The reduced angle set (fewer azimuths near the poles) still integrates
    the m = 0 terms exactly."""
    w = S.lognormal_blob(n=8, nmu=8, nphi=16, nleg=64)
    rte = S.make_rte(w)
    y = rte.plan_array("ymat").astype(np.float64)
    wm = rte.plan_array("wmat").astype(np.float64)
    g = wm.T @ y
    # m(J) = 0 columns: index J = l(l+1) for l <= MM (ML = 7, MM = 7 here)
    j0 = [l * (l + 1) for l in range(8)]
    sub = g[np.ix_(j0, j0)]
    assert np.max(np.abs(sub - np.eye(len(j0)))) < 2e-6
