"""The oracle against original SHDOM: the reference's own verification of its
polarized solver (tests/test_shdom.py:175-277), on SHDOM's output files.  The
property file is read by at3d_amd.propfile (READ_PROPERTIES' 'P' format); the
oracle solves with adaptive splitting and INTERPMETHOD 'OO' and renders
without CORRECTINTERPOLATE.  Pinned: the iteration count and grid size of
SHDOM's run, its source function to the reference test's 5e-7, and its I, Q,
U radiances to the reference test's tolerances."""
import numpy as np
import pytest

from at3d_amd.propfile import read_property_file
from tests import shdom_rico as R

ref = pytest.importorskip("oracle.ref_shdom")
pytestmark = pytest.mark.skipif(not ref.available(), reason="oracle library not built")


def test_property_file_reader():
    pf = read_property_file(R.PRP, nleg=8)
    assert (pf.proptype, pf.npx, pf.npy, pf.npz, pf.numphase) == ("P", 32, 36, 27, 18)
    assert pf.delx == np.float32(0.02) and pf.dely == np.float32(0.02)
    assert pf.zlevels[0] == 0.0 and pf.zlevels[1] == np.float32(0.44) and pf.zlevels[-1] == np.float32(1.44)
    # the first table (Rayleigh-like) as the file writes it: a1 = 1, 0, 0.47945
    assert np.array_equal(pf.legenp[0, :3, 0], np.float32([1.0, 0.0, 0.47945]))
    assert pf.legenp[3, 1, 0] == np.float32(1.39723)
    assert pf.nleg == pf.legenp.shape[1] - 1 >= 877
    assert pf.iphasep.min() == 1 and pf.iphasep.max() == 18
    assert np.all(pf.albedop >= 0) and np.all(pf.albedop <= 1)


@pytest.fixture(scope="module")
def solved():
    pf = read_property_file(R.PRP, nleg=8)
    sp = ref.Species(pf.extinctp, pf.albedop, pf.legenp, pf.iphasep)
    c = R.CONFIG
    sc = ref.Scene(delx=pf.delx, dely=pf.dely, zlevels=pf.zlevels, species=[sp], solarmu=-0.5, solaraz=0.0,
                   gndalbedo=0.05, wavelength=0.672, num_mu_bins=c["num_mu_bins"], num_phi_bins=c["num_phi_bins"],
                   bc="open", shacc=c["spherical_harmonics_accuracy"], solacc=c["solution_accuracy"],
                   splitacc=c["split_accuracy"], tautol=c["tautol"], nstokes=3, interpmethod="OO",
                   correctinterpolate=False)
    return ref.solve(sc, maxiter=100)


def test_oracle_source_matches_shdom(solved):
    st = solved
    assert st.iters == R.ITERATIONS and st.npts == R.NPTS
    truth = R.source_truth()
    got = st.source[:, :st.npts]
    assert got.shape == truth.shape
    assert np.allclose(got, truth, atol=R.SOURCE_ATOL)


def test_oracle_radiances_match_shdom(solved):
    x, y, z, mu, phi = R.rays()
    stokes = ref.render(solved, x, y, z, mu, phi)
    rad = R.radiance_truth()
    assert rad.shape == (x.size, 5)
    for k in range(3):
        assert np.allclose(stokes[k], rad[:, 2 + k], atol=R.STOKES_ATOL[k]), k
    # much closer than the reference test asks: same algorithm, same precision
    assert np.max(np.abs(stokes[0] - rad[:, 2])) < 1e-4
