"""at3d_amd.propfile against the reference's own READ_PROPERTY_SIZE and
READ_PROPERTIES (shdomsub3.f:286-560, compiled into the oracle library) on
the same files: the polarized ('P') RICO property file of the reference's
tests and a tabulated ('T') file written here (tables of different lengths,
records spread over several lines, one point outside the grid)."""
import gzip
import shutil

import numpy as np
import pytest

from at3d_amd.propfile import read_property_file
from tests import shdom_rico as R

ref = pytest.importorskip("oracle.ref_shdom")
pytestmark = pytest.mark.skipif(not ref.available(), reason="oracle library not built")


def _reference_read(path, nstleg, nleg):
    p = ref.s(str(path), 80)
    nleg_io, nstleg_io = ref.i4(nleg), ref.i4(nstleg)
    npx, npy, npz, numphase, maxleg, maxpgl = ref.i4(), ref.i4(), ref.i4(), ref.i4(), ref.i4(), ref.i4()
    delx, dely = ref.r4(), ref.r4()
    ref.fcall("READ_PROPERTY_SIZE", p, nstleg_io, nleg_io, npx, npy, npz, numphase, maxleg, maxpgl, delx, dely)
    n = npx.value * npy.value * npz.value
    zl = np.zeros(npz.value, np.float32)
    tempp, extp, albp = (np.zeros(n, np.float32) for _ in range(3))
    legenp = np.zeros(maxpgl.value, np.float32)
    iphasep = np.zeros(n, np.int32)
    proptype, maxasym = ref.s("", 1), ref.r4()
    nleg2 = ref.i4(nleg)
    npx2, npy2, npz2, nph2 = ref.i4(npx.value), ref.i4(npy.value), ref.i4(npz.value), ref.i4()
    ref.fcall("READ_PROPERTIES", ref.s(str(path), 80), npx2, npy2, npz2, ref.i4(nstleg), ref.i4(maxleg.value),
              nleg2, ref.i4(npz.value), ref.i4(n), ref.i4(maxpgl.value), ref.lg(True), proptype, delx, dely, zl,
              maxasym, tempp, extp, albp, legenp, nph2, iphasep)
    shape = (npx.value, npy.value, npz.value)
    return dict(nleg=nleg2.value, numphase=nph2.value, delx=delx.value, dely=dely.value, zlevels=zl,
                tempp=tempp.reshape(shape), ext=extp.reshape(shape), alb=albp.reshape(shape),
                iph=iphasep.reshape(shape), legenp=legenp, maxasym=maxasym.value)


def _compare(pf, want, nstleg):
    assert pf.nleg == want["nleg"] and pf.numphase == want["numphase"]
    assert pf.delx == want["delx"] and pf.dely == want["dely"]
    assert np.array_equal(pf.zlevels, want["zlevels"])
    for got, key in ((pf.tempp, "tempp"), (pf.extinctp, "ext"), (pf.albedop, "alb"), (pf.iphasep, "iph")):
        assert np.array_equal(got, want[key]), key
    n = want["numphase"] * (want["nleg"] + 1) * nstleg
    leg = want["legenp"][:n].reshape(want["numphase"], want["nleg"] + 1, nstleg)   # LEGENP(NSTLEG, 0:NLEG, NUMPHASE)
    if nstleg == 6:
        assert np.array_equal(pf.legenp, np.transpose(leg, (2, 1, 0)))
    else:
        assert np.array_equal(pf.legenp, leg[:, :, 0].T)
    assert np.float32(pf.maxasym) == np.float32(want["maxasym"])


def test_polarized_property_file_matches_read_properties(tmp_path):
    path = tmp_path / "rico.prp"
    with gzip.open(R.PRP, "rb") as f, open(path, "wb") as g:
        shutil.copyfileobj(f, g)
    if len(str(path)) > 80:
        pytest.skip("temporary path longer than the reference's CHARACTER*80")
    _compare(read_property_file(str(path), nleg=8), _reference_read(path, 6, 8), 6)


def test_tabulated_property_file_matches_read_properties(tmp_path):
    path = tmp_path / "t.prp"
    rng = np.random.default_rng(4)
    lines = ["Tabulated phase function property file", "3 2 4", "0.05 0.04 0.0 0.5 1.0 1.5", "2"]
    chi1 = np.round(rng.uniform(0, 2, 5), 5)
    chi2 = np.round(rng.uniform(0, 2, 11), 5)
    lines.append("5 " + " ".join(f"{v:.5f}" for v in chi1))
    # a table spread over three lines
    lines.append("11 " + " ".join(f"{v:.5f}" for v in chi2[:4]))
    lines.append(" ".join(f"{v:.5f}" for v in chi2[4:8]))
    lines.append(" ".join(f"{v:.5f}" for v in chi2[8:]))
    for ix in range(1, 4):
        for iy in range(1, 3):
            for iz in range(1, 5):
                ext, alb = rng.uniform(0, 20), rng.uniform(0.5, 1.0)
                lines.append(f"{ix} {iy} {iz} 280.0 {ext:.4f} {alb:.5f} {1 + (ix + iz) % 2}")
    lines.append("9 1 1 280.0 5.0 0.9 1")   # outside the grid: skipped
    path.write_text("\n".join(lines) + "\n")
    if len(str(path)) > 80:
        pytest.skip("temporary path longer than the reference's CHARACTER*80")
    _compare(read_property_file(str(path), nleg=8), _reference_read(path, 1, 8), 1)


def _write(path, lines):
    path.write_text("\n".join(lines) + "\n")
    if len(str(path)) > 80:
        pytest.skip("temporary path longer than the reference's CHARACTER*80")


def _compare_loose(pf, want):
    """'E' and standard formats: grid arrays and the tables as the reference
    leaves them."""
    assert pf.nleg == want["nleg"]
    assert pf.delx == want["delx"] and pf.dely == want["dely"]
    assert np.array_equal(pf.zlevels, want["zlevels"])
    for got, key in ((pf.tempp, "tempp"), (pf.extinctp, "ext"), (pf.albedop, "alb")):
        assert np.array_equal(got, want[key]), key
    assert np.float32(pf.maxasym) == np.float32(want["maxasym"])


@pytest.mark.parametrize("numl", [3, 12])
def test_extinction_only_property_file_matches_read_properties(tmp_path, numl):
    rng = np.random.default_rng(numl)
    chi = np.round(rng.uniform(0, 2, numl), 5)
    lines = ["E", "3 2 4", "0.05 0.04 0.0 0.5 1.0 1.5", "290.0 280.0 270.0 260.0",
             "0.99 {} ".format(numl) + " ".join(f"{v:.5f}" for v in chi)]
    for ix in range(1, 4):
        for iy in range(1, 3):
            for iz in range(1, 5):
                lines.append(f"{ix} {iy} {iz} {rng.uniform(0, 20):.4f}")
    path = tmp_path / "e.prp"
    _write(path, lines)
    want = _reference_read(path, 1, 8)
    pf = read_property_file(str(path), nleg=8)
    _compare_loose(pf, want)
    assert np.array_equal(pf.legenp[:, 0], want["legenp"][:pf.nleg + 1])
    assert np.all(pf.iphasep == 1)


def test_standard_property_file_matches_read_properties(tmp_path):
    rng = np.random.default_rng(7)
    lines = ["2 2 3", "0.05 0.05 0.0 0.5 1.0"]
    numls = {}
    for ix in range(1, 3):
        for iy in range(1, 3):
            for iz in range(1, 4):
                numl = int(rng.integers(2, 12))
                chi = np.round(rng.uniform(0, 2, numl), 5)
                numls[(ix, iy, iz)] = numl
                lines.append(f"{ix} {iy} {iz} 280.0 {rng.uniform(0, 20):.4f} {rng.uniform(0.5, 1):.5f} {numl} "
                             + " ".join(f"{v:.5f}" for v in chi))
    path = tmp_path / "s.prp"
    _write(path, lines)
    want = _reference_read(path, 1, 8)
    pf = read_property_file(str(path), nleg=8)
    _compare_loose(pf, want)
    n = pf.npx * pf.npy * pf.npz
    ref_leg = want["legenp"][:n * (want["nleg"] + 1)].reshape(n, want["nleg"] + 1)
    assert np.array_equal(pf.legenp.T, ref_leg)
