"""The polarized (NSTOKES = 3) host plan against the reference's own routines.

* The dense six-element transform tables, applied to random source / radiance
  terms, give what SH_TO_DO and DO_TO_SH (shdomsub1.f:2789-3160) give on the
  same terms, for every zenith angle, FFT and DFT azimuth paths alike, at L = 7
  and L = 15.
* The sun's polarized harmonics equal INIT_SOLUTION's YLMSUN (YLMALL with
  NSTLEG = 6); the I-I and I-Q phase tables equal PRECOMPUTE_PHASE_CHECK's.
* The prepared a1, a2, a3 and b1 tables equal the reference's LEGEN after
  TRANSFER_PA_TO_GRID and PREPARE_PROP (delta-M f subtracted from a1 to a4).
"""
import numpy as np
import pytest

from at3d_amd import synthetic as S
from tests.scenes import pol_blob

ref = pytest.importorskip("oracle.ref_shdom")
pytestmark = pytest.mark.skipif(not ref.available(), reason="oracle library not built")


def _pair(nmu, nphi):
    w = pol_blob(nmu=nmu, nphi=nphi)
    rte = S.make_rte(w)
    st = ref.setup(ref.scene_from_workload(w))
    ref.init_solution(st)
    return w, rte, st


@pytest.mark.parametrize("nmu,nphi", [(8, 16), (16, 32)])
def test_transform_tables_match_sh_to_do_and_do_to_sh(nmu, nphi):
    w, rte, st = _pair(nmu, nphi)
    d = st.d
    nlm, ml, mm, nmu_, nphi0max = d["nlm"], d["ml"], d["mm"], d["nmu"], d["nphi0max"]
    Y = rte.plan_array("ypol")
    Wt = rte.plan_array("wpol")
    nphi0 = np.asarray(d["nphi0"])
    first = np.concatenate([[0], np.cumsum(nphi0)])
    rng = np.random.default_rng(5)
    npts = 3
    sh = rng.standard_normal((3, npts, nlm)).astype(np.float32)
    indata = ref.farr(np.transpose(sh, (0, 1, 2)).reshape(3, npts * nlm), np.float32)
    shptr = np.arange(npts + 1, dtype=np.int32) * nlm
    for imu in range(nmu_):
        out = np.zeros((3, nphi0max, npts), np.float32, order="F")
        ref.fcall("SH_TO_DO", ref.i4(npts), ref.i4(3), ref.i4(6), ref.i4(ml), ref.i4(mm), ref.i4(nlm),
                  ref.i4(nmu_), ref.i4(nphi0max), ref.i4(nphi0[imu]), ref.i4(imu + 1), d["fftflag"], d["cmu1"],
                  d["cphi1"], d["wphisave"], shptr, indata, out)
        k = slice(first[imu], first[imu + 1])
        got_i = np.einsum("kj,pj->kp", Y[0, k], sh[0])
        got_q = np.einsum("kj,pj->kp", Y[1, k], sh[1]) + np.einsum("kj,pj->kp", Y[4, k], sh[2])
        got_u = np.einsum("kj,pj->kp", Y[5, k], sh[1]) + np.einsum("kj,pj->kp", Y[2, k], sh[2])
        n0 = nphi0[imu]
        scale = np.abs(out).max()
        for got, want in ((got_i, out[0, :n0]), (got_q, out[1, :n0]), (got_u, out[2, :n0])):
            assert np.allclose(got, want, rtol=0, atol=2e-5 * scale), imu
    # the analysis: running sums over the zenith angles
    do = rng.standard_normal((3, nphi0max, npts, nmu_)).astype(np.float32)
    acc = np.zeros((3, npts * nlm), np.float32, order="F")
    for imu in range(nmu_):
        ref.fcall("DO_TO_SH", ref.i4(npts), ref.i4(3), ref.i4(6), ref.i4(ml), ref.i4(mm), ref.i4(nlm),
                  ref.i4(nmu_), ref.i4(nphi0max), ref.i4(nphi0[imu]), ref.i4(imu + 1), d["fftflag"], d["cmu2"],
                  d["cphi2"], d["wphisave"], shptr, ref.farr(do[..., imu], np.float32), acc)
    want = acc.reshape(3, npts, nlm)
    q = np.concatenate([do[0, :nphi0[i], :, i] for i in range(nmu_)])
    qq = np.concatenate([do[1, :nphi0[i], :, i] for i in range(nmu_)])
    uu = np.concatenate([do[2, :nphi0[i], :, i] for i in range(nmu_)])
    got_i = np.einsum("kj,kp->pj", Wt[0], q)
    got_q = np.einsum("kj,kp->pj", Wt[1], qq) + np.einsum("kj,kp->pj", Wt[4], uu)
    got_u = np.einsum("kj,kp->pj", Wt[5], qq) + np.einsum("kj,kp->pj", Wt[2], uu)
    scale = np.abs(want).max()
    for got, wnt in ((got_i, want[0]), (got_q, want[1]), (got_u, want[2])):
        assert np.allclose(got, wnt, rtol=0, atol=2e-5 * scale)


@pytest.mark.parametrize("nmu,nphi", [(8, 16), (16, 32)])
def test_sun_harmonics_and_phase_tables(nmu, nphi):
    w, rte, st = _pair(nmu, nphi)
    ysun = rte.plan_array("ylmsun")
    assert np.allclose(ysun, st.ylmsun[0], rtol=1e-5, atol=1e-7)
    pt = ref.precompute_phase(st)
    assert np.allclose(rte.plan_array("phasetab"), pt[0], rtol=1e-5, atol=1e-7)
    assert np.allclose(rte.plan_array("phasetab2"), pt[1], rtol=1e-5, atol=1e-7)
    # the I-Q table has signal: the Rayleigh species polarizes
    assert np.abs(pt[1]).max() > 1e-3


def test_prepared_phase_matrix_elements():
    w, rte, st = _pair(8, 16)
    legen = st.legen   # (6, nleg+1, numphase) after PREPARE_PROP
    for name, e in (("legen", 0), ("legen2", 1), ("legen3", 2), ("legen5", 4)):
        assert np.array_equal(rte.plan_array(name), legen[e].T), name


def test_polarized_options_checked():
    w = pol_blob()
    with pytest.raises(ValueError):
        S.make_rte(S.Workload("v", w.delx, w.dely, w.z, w.species, bc=w.bc, nstokes=4))
