"""Sensor geometry (at3d/sensor.py) pinned by geometry and quadrature facts:
the reference's rendered-sensor fixtures are netCDF4/HDF5 files this image
cannot read, so the projections are checked against what they must satisfy.

* orthographic: parallel rays, one pixel every resolution step from the low
  corner of the box's footprint on the sensor plane, one step past the high
  corner (np.arange(x_s, x_e + res, res) in the reference); every box corner
  projects inside the pixel span; sub-pixel rays stay within half a pixel;
* perspective: the central ray looks along lookat - position, the outermost
  pixel centres sit half a pixel inside the field of view, the up vector sets
  the image's y axis;
* sub-pixel generators: Gauss-Legendre nodes and normalised weights
  (gaussian), weights summing to one per pixel; average_subpixel_rays is the
  weighted sum (AVERAGE_SUBPIXEL_RAYS, util.f90:484).
"""
import numpy as np
import pytest

from at3d_amd import sensor as SN

BOX = (0.0, 0.0, 0.0, 0.64, 0.48, 0.5)


def _back_to(z, s):
    """Where each ray's photon was at height z (tracing back from the sensor)."""
    mu = np.asarray(s["ray_mu"])
    phi = np.asarray(s["ray_phi"])
    t = np.sqrt(1 - mu ** 2) / mu
    dz = np.asarray(s["ray_z"]) - z
    # the photon travels along -(direction towards the sensor)
    return np.asarray(s["ray_x"]) - dz * t * np.cos(phi), np.asarray(s["ray_y"]) - dz * t * np.sin(phi)


@pytest.mark.parametrize("zen,az", [(0.0, 0.0), (45.0, 30.0), (60.0, 200.0), (26.1, 95.0)])
def test_orthographic_footprint(zen, az):
    res = 0.02
    s = SN.orthographic_projection(0.86, BOX, res, res, az, zen)
    nx, ny = s["image_shape"]
    assert s["ray_x"].size == nx * ny and np.all(s["ray_z"] == BOX[5])
    assert np.allclose(s["ray_mu"], np.cos(np.deg2rad(zen))) and np.allclose(s["ray_phi"], np.deg2rad(az))
    x = s["ray_x"].reshape(ny, nx)
    y = s["ray_y"].reshape(ny, nx)
    assert np.allclose(np.diff(x, axis=1), res) and np.allclose(np.diff(y, axis=0), res)
    # every corner of the box, followed along the view direction up to the
    # sensor plane, lands within the pixel span; the span starts at the
    # lowest such landing point and ends at most one step past the highest
    mu, phi = np.cos(np.deg2rad(zen)), np.deg2rad(az)
    t = np.sqrt(1 - mu ** 2) / mu
    land_x, land_y = [], []
    for cx in (BOX[0], BOX[3]):
        for cy in (BOX[1], BOX[4]):
            for cz in (BOX[2], BOX[5]):
                land_x.append(cx + (BOX[5] - cz) * t * np.cos(phi))
                land_y.append(cy + (BOX[5] - cz) * t * np.sin(phi))
    assert x.min() == pytest.approx(min(land_x), abs=1e-12)
    assert y.min() == pytest.approx(min(land_y), abs=1e-12)
    assert max(land_x) <= x.max() + 1e-12 and x.max() < max(land_x) + 2 * res + 1e-12
    assert max(land_y) <= y.max() + 1e-12 and y.max() < max(land_y) + 2 * res + 1e-12
    # the photons of the first column and row pass the box's low edge at some height
    bx, _ = _back_to(BOX[2], s)
    assert np.all(np.isfinite(bx))


@pytest.mark.parametrize("method,args,nrays", [(SN.gaussian, {"degree": (2, 3)}, 6),
                                               (SN.uniform, {"nrays": (3, 2)}, 6),
                                               (SN.stochastic, {"nrays": (2, 2), "seed": (1, 2)}, 4)])
def test_orthographic_subpixel_rays(method, args, nrays):
    res = 0.04
    base = SN.orthographic_projection(0.86, BOX, res, res, 30.0, 40.0)
    s = SN.orthographic_projection(0.86, BOX, res, res, 30.0, 40.0,
                                   sub_pixel_ray_args=dict(method=method, **args))
    npix = base["cam_x"].size
    assert s["ray_x"].size == npix * nrays
    assert np.array_equal(SN.rays_per_pixel(s), np.full(npix, nrays))
    w = np.bincount(s["pixel_index"], weights=s["ray_weight"], minlength=npix)
    assert np.allclose(w, 1.0)
    px = np.asarray(base["ray_x"])[s["pixel_index"]]
    py = np.asarray(base["ray_y"])[s["pixel_index"]]
    assert np.all(np.abs(s["ray_x"] - px) <= res / 2 + 1e-12)
    assert np.all(np.abs(s["ray_y"] - py) <= res / 2 + 1e-12)
    # a constant averages to itself, a linear field to its pixel-centre value
    # for the symmetric rules
    assert np.allclose(SN.average_subpixel_rays(s, np.full(s["ray_x"].size, 2.5)), 2.5)
    if method is not SN.stochastic:
        got = SN.average_subpixel_rays(s, 3.0 * s["ray_x"] - 2.0 * s["ray_y"])
        assert np.allclose(got, 3.0 * base["ray_x"] - 2.0 * base["ray_y"])


def test_gaussian_rule_integrates_polynomials():
    """Gauss-Legendre with n nodes is exact to degree 2n-1 over a pixel."""
    off, w = SN.gaussian(5, 3)
    assert np.allclose(w.sum(axis=1), 1.0)
    for k in range(6):
        exact = (1 - (-1) ** (k + 1)) / (k + 1) / 2.0      # mean of x^k over [-1, 1]
        assert np.allclose((w * off ** k).sum(axis=1), exact)


def test_perspective_geometry():
    pos, look, up = (0.3, 0.2, 5.0), (0.3, 0.2, 0.0), (0.0, 1.0, 0.0)
    nx, ny, fov = 9, 7, 20.0
    s = SN.perspective_projection(0.86, fov, nx, ny, pos, look, up)
    assert s["ray_x"].size == nx * ny
    assert np.all(s["ray_x"] == np.float32(0.3)) and np.all(s["ray_z"] == np.float32(5.0))
    mu = s["ray_mu"].reshape(ny, nx)
    phi = s["ray_phi"].reshape(ny, nx)
    # the central pixel looks straight down: its photons travel straight up
    assert mu[ny // 2, nx // 2] == pytest.approx(1.0, abs=1e-12)
    # pixel centres along the longer (x) axis sit half a pixel inside the fov:
    # the half angle of the outermost column is atan((1 - 1/nx) tan(fov/2))
    # (to single precision: the camera axis and 1/f are REAL in the reference)
    want = np.arctan((1 - 1.0 / nx) * np.tan(np.deg2rad(fov) / 2))
    assert np.arccos(mu[ny // 2, 0]) == pytest.approx(want, rel=1e-6)
    assert np.arccos(mu[ny // 2, -1]) == pytest.approx(want, rel=1e-6)
    # along y the image is ny/nx as tall
    want_y = np.arctan((ny / nx) * (1 - 1.0 / ny) * np.tan(np.deg2rad(fov) / 2))
    assert np.arccos(mu[0, nx // 2]) == pytest.approx(want_y, rel=1e-6)
    # symmetric image: mirrored pixels have the same zenith
    assert np.allclose(mu, mu[::-1, ::-1], rtol=1e-6)
    # image y follows the up vector: the top row looks towards +y, so its
    # photons arrive from -y (direction towards the sensor has azimuth -90 deg)
    d_top = np.array([np.cos(phi[-1, nx // 2]), np.sin(phi[-1, nx // 2])])
    assert d_top[1] < 0 and abs(d_top[0]) < 1e-9


def test_perspective_subpixel_rays_fill_each_pixel():
    s = SN.perspective_projection(0.86, 10.0, 4, 4, (0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0),
                                  sub_pixel_ray_args={"method": SN.gaussian, "degree": (2, 2)})
    assert np.array_equal(SN.rays_per_pixel(s), np.full(16, 4))
    assert np.allclose(np.bincount(s["pixel_index"], weights=s["ray_weight"]), 1.0)
    assert np.all(s["ray_mu"] > np.cos(np.deg2rad(10.0)))     # every ray inside the field of view
