"""The oracle's polarized (NSTOKES = 3) path pinned by what the physics fixes.

* A phase matrix whose only nonzero element is a1 scatters no polarization:
  Q and U stay zero, and I, the fluxes and the rendered radiances equal the
  scalar solve's (the polarized path represents I in another real basis for
  m != 0, so the equality holds to rounding).
* An optically thin Rayleigh layer over a black surface, seen from above, is
  single scattering: the degree of linear polarization of every view is
  (1 - cos^2 T) / (1 + cos^2 T) at scattering angle T, the light is
  polarized perpendicular to the scattering plane (Q < 0, U = 0 in the
  principal plane), with and without delta-M (the TMS single-scatter table
  with its rotation into the meridian plane, or the spherical-harmonic
  source of the l <= 2 Rayleigh matrix).
"""
import numpy as np
import pytest

from at3d_amd import synthetic as S

ref = pytest.importorskip("oracle.ref_shdom")
pytestmark = pytest.mark.skipif(not ref.available(), reason="oracle library not built")


def _a1_only(table):
    t = np.zeros((6,) + table.shape)
    t[0] = table
    return t


def test_a1_only_phase_matrix_matches_scalar():
    w = S.lognormal_blob(n=8, radius=0.15, seed=0, nleg=64, bc="open", nmu=8, nphi=16)
    e, a, leg = w.species[0]
    scal = ref.solve(ref.scene_from_workload(w), maxiter=100)
    w3 = S.Workload(w.name, w.delx, w.dely, w.z, [(e, a, _a1_only(leg))], bc=w.bc, nmu=w.nmu,
                    nphi=w.nphi, solacc=w.solacc, nstokes=3)
    pol = ref.solve(ref.scene_from_workload(w3), maxiter=100)
    assert pol.iters == scal.iters
    n = scal.npts
    assert np.allclose(pol.fluxes[:, :n], scal.fluxes[:, :n], rtol=1e-5, atol=1e-7)
    for k in (1, 2):
        assert np.all(pol.source[k] == 0.0) and np.all(pol.radiance[k] == 0.0)
    rng = np.random.default_rng(3)
    m = 40
    x = rng.uniform(0, 7 * w.delx, m)
    y = rng.uniform(0, 7 * w.dely, m)
    z = np.full(m, float(w.z[-1]))
    mu = rng.uniform(0.2, 1.0, m)
    phi = rng.uniform(0, 2 * np.pi, m)
    i1 = ref.render(scal, x, y, z, mu, phi)
    s3 = ref.render(pol, x, y, z, mu, phi)
    assert np.allclose(s3[0], i1, rtol=2e-5, atol=1e-7)
    assert np.all(s3[1:] == 0.0)


@pytest.mark.parametrize("deltam", [True, False])
def test_thin_rayleigh_single_scatter_polarization(deltam):
    n, nz = 4, 3
    z = np.array([0.0, 0.5, 1.0], np.float32)
    tau = 0.002
    ext = np.full((n, n, nz), tau / 1.0, np.float32)
    alb = np.ones((n, n, nz), np.float32)
    w = S.Workload("rayleigh", 0.1, 0.1, z, [(ext, alb, S.rayleigh_phase_matrix(0.0))], bc="periodic",
                   nmu=8, nphi=16, solacc=1e-6, gndalbedo=0.0, solarmu=-0.5, deltam=deltam, nstokes=3)
    st = ref.solve(ref.scene_from_workload(w), maxiter=50)
    zen = np.deg2rad(np.array([0.0, 20.0, 40.0, 60.0, 30.0, 50.0, 45.0]))
    az = np.deg2rad(np.array([0.0, 0.0, 0.0, 180.0, 90.0, 135.0, 250.0]))
    mu = np.cos(zen)
    x = np.full(mu.size, 0.15)
    y = np.full(mu.size, 0.15)
    zz = np.full(mu.size, 1.0)
    I, Q, U = ref.render(st, x, y, zz, mu, az)
    # scattering angle between the sun's propagation and the view's
    mu0, phi0 = -0.5, 0.0
    cos_t = np.sqrt(1 - mu0 ** 2) * np.sqrt(1 - mu ** 2) * np.cos(az - phi0) + mu0 * mu
    dolp = (1 - cos_t ** 2) / (1 + cos_t ** 2)
    got = np.sqrt(Q ** 2 + U ** 2) / I
    assert np.allclose(got, dolp, atol=5e-3)
    # principal plane (azimuth 0 or 180 deg): polarized perpendicular to it
    # (the view at 60 deg, azimuth 180 deg, is exact backscatter: unpolarized)
    pp = np.isclose(np.sin(az), 0.0)
    assert np.all(Q[pp & (dolp > 0.01)] < 0) and np.all(np.abs(U[pp]) < 1e-3 * I[pp])
    # radiance: F0 / |mu0| (flux on a horizontal surface) x omega tau / mu x P11 / (4 pi)
    p11 = 0.75 * (1 + cos_t ** 2)
    i_ss = tau * p11 / (4 * np.pi * mu) / abs(mu0)
    assert np.allclose(I, i_ss, rtol=3e-2)
