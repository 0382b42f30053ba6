"""Sensor geometry as plain dicts of numpy arrays (the fields of the
reference's sensor xr.Dataset, at3d/sensor.py:31-107).

A sensor holds per-pixel 'cam_*' arrays and per-ray 'ray_*' arrays, with
'ray_mu' the cosine zenith and 'ray_phi' the azimuth (radians) of the
direction TOWARDS the sensor, 'pixel_index' (the pixel each ray belongs to)
and 'ray_weight' (the ray's weight in its pixel; the weights of one pixel sum
to one).  Without sub-pixel rays each pixel is one ray of weight one.
"""
from __future__ import annotations

import numpy as np


def make_sensor_dataset(x, y, z, mu, phi, stokes, wavelength):
    """Per-pixel geometry with one ray per pixel (at3d/sensor.py:31-107 with
    fill_ray_variables=True)."""
    x, y, z, mu, phi = (np.asarray(v) for v in (x, y, z, mu, phi))
    for v, name in zip((x, y, z, mu, phi), ("x", "y", "z", "mu", "phi")):
        if v.ndim != 1:
            raise ValueError("'{}' should be a 1-D np.ndarray".format(name))
    if not all(x.size == v.size for v in (y, z, mu, phi)):
        raise ValueError("All of x, y, z, mu, phi should have the same size.")
    if not np.all(z >= 0.0):
        raise ValueError("All altitudes (z) must be >= 0.0")
    if np.any(mu == 0.0):
        raise ValueError("'mu' values of 0.0 are not allowed.")
    stokes = [str(v) for v in np.atleast_1d(stokes)]
    if not stokes or any(v not in ("I", "Q", "U") for v in stokes) or len(set(stokes)) != len(stokes):
        raise ValueError("stokes must name distinct components among 'I', 'Q', 'U' (NSTOKES <= 3)")
    s = {"wavelength": float(wavelength), "stokes": stokes,
         "cam_x": x.astype(np.float64), "cam_y": y.astype(np.float64),
         "cam_z": z.astype(np.float64), "cam_mu": mu.astype(np.float64),
         "cam_phi": phi.astype(np.float64)}
    for k in ("x", "y", "z", "mu", "phi"):
        s["ray_" + k] = s["cam_" + k].copy()
    s["pixel_index"] = np.arange(x.size)
    s["ray_weight"] = np.ones(x.size)
    s["use_subpixel_rays"] = False
    return s


# ----------------------------------------------------------------------------
# sub-pixel ray generators: (position offsets in units of half a pixel, weights)
# per pixel along one image axis (at3d/sensor.py:109-186)
# ----------------------------------------------------------------------------
def stochastic(npixels, nrays, seed=None):
    """`nrays` offsets per pixel drawn uniformly from [-1, 1) and equal
    weights 1/nrays (at3d/sensor.py:109)."""
    rng = np.random.RandomState(seed) if seed is not None else np.random
    offsets = rng.uniform(-1.0, 1.0, size=(npixels, nrays))
    return offsets, np.full((npixels, nrays), 1.0 / nrays)


def uniform(npixels, nrays):
    """`nrays` evenly spaced offsets centred in equal sub-intervals of the
    pixel, equal weights (at3d/sensor.py:146)."""
    half = 1.0 / nrays
    offsets = np.linspace(-1.0 + half, 1.0 - half, nrays)
    return offsets, np.full((npixels, nrays), 1.0 / nrays)


def gaussian(npixels, degree):
    """Gauss-Legendre nodes of `degree` points on [-1, 1] as offsets, with the
    quadrature weights normalised to one (at3d/sensor.py:152)."""
    nodes, weights = np.polynomial.legendre.leggauss(degree)
    offsets = np.broadcast_to(nodes, (npixels, degree)).copy()
    w = np.broadcast_to(weights / weights.sum(), (npixels, degree)).copy()
    return offsets, w


def _split_ray_args(args):
    """The generator and its per-axis keyword arguments: every argument but
    'method' holds one value for the image x axis and one for the y axis."""
    method = args["method"]
    kx, ky = {}, {}
    for k, v in args.items():
        if k == "method":
            continue
        if np.size(v) != 2:
            raise ValueError(f"sub-pixel ray argument '{k}' needs one value per image axis")
        kx[k], ky[k] = v[0], v[1]
    return method, kx, ky


def _subpixel(npx_x, npx_y, args):
    """Offsets (image-axis units of half a pixel) and weights of every ray,
    pixel by pixel in the pixel order (x fastest): each pixel takes the outer
    product of its x and y rays, the y ray index varying fastest."""
    method, kx, ky = _split_ray_args(args)
    npix = npx_x * npx_y
    ox, wx = method(npix, **kx)
    oy, wy = method(npix, **ky)
    ox, wx, oy, wy = (np.asarray(a, np.float64) for a in (ox, wx, oy, wy))
    if ox.ndim == 1:
        ox = np.broadcast_to(ox, (npix, ox.size))
    if oy.ndim == 1:
        oy = np.broadcast_to(oy, (npix, oy.size))
    nrx, nry = ox.shape[1], oy.shape[1]
    off_x = np.repeat(ox[:, :, None], nry, axis=2)
    off_y = np.repeat(oy[:, None, :], nrx, axis=1)
    weight = wx[:, :, None] * wy[:, None, :]
    return off_x.reshape(npix, -1), off_y.reshape(npix, -1), weight.reshape(npix, -1)


def _attach_rays(s, x, y, z, mu, phi, weight):
    """Store per-ray arrays of shape (npix, rays per pixel) on sensor `s`."""
    npix, k = weight.shape
    s["ray_x"], s["ray_y"], s["ray_z"] = (np.asarray(a, np.float64).ravel() for a in (x, y, z))
    s["ray_mu"], s["ray_phi"] = (np.asarray(a, np.float64).ravel() for a in (mu, phi))
    s["pixel_index"] = np.repeat(np.arange(npix), k)
    s["ray_weight"] = weight.ravel()
    s["use_subpixel_rays"] = True
    return s


def rays_per_pixel(sensor):
    """Number of rays of each pixel (pixels in order), from 'pixel_index'."""
    return np.bincount(np.asarray(sensor["pixel_index"]),
                       minlength=np.size(sensor["cam_x"])).astype(np.int32)


def average_subpixel_rays(sensor, values):
    """Pixel values from ray values: the ray-weighted sum over each pixel's
    rays (AVERAGE_SUBPIXEL_RAYS, src/util.f90:484)."""
    v = np.asarray(values, np.float64)
    w = np.asarray(sensor["ray_weight"], np.float64)
    idx = np.asarray(sensor["pixel_index"])
    npix = np.size(sensor["cam_x"])
    if v.ndim == 1:
        return np.bincount(idx, weights=v * w, minlength=npix)
    return np.stack([np.bincount(idx, weights=row * w, minlength=npix) for row in v])


# ----------------------------------------------------------------------------
# projections
# ----------------------------------------------------------------------------
def orthographic_projection(wavelength, bounding_box, x_resolution, y_resolution, azimuth, zenith,
                            altitude="TOA", stokes="I", sub_pixel_ray_args=None):
    """Parallel rays viewing ``bounding_box`` = (xmin, ymin, zmin, xmax, ymax,
    zmax) at azimuth/zenith (degrees, the direction of the photons), as
    at3d/sensor.py:188-347.

    Geometry: every ray leaves a sensor plane at ``altitude`` (the box top
    for 'TOA').  Following a photon back from the plane down to height z
    moves it horizontally by (altitude - z) tan(zenith) along the azimuth, so
    the box's footprint on the plane spans, in x, from xmin plus the smaller
    of that shift at zmin and at zmax to xmax plus the larger (likewise in y).
    Pixels sit on that footprint every x_resolution / y_resolution km from
    its low corner, and one more step is taken past the high corner so that
    a pixel lies on or beyond the domain edge.  Pixels are ordered with x
    fastest.  ``sub_pixel_ray_args`` {'method': stochastic | uniform |
    gaussian, <argument>: (x value, y value)} spreads each pixel's rays over
    +-half a pixel in the plane."""
    mu = float(np.cos(np.deg2rad(zenith)))
    phi = float(np.deg2rad(azimuth))
    xmin, ymin, zmin, xmax, ymax, zmax = (float(v) for v in bounding_box)
    alt = zmax if altitude == "TOA" else float(altitude)
    tan_zen = np.sqrt(1.0 - mu * mu) / mu
    shift_x = tan_zen * np.cos(phi) * np.array([alt - zmin, alt - zmax])
    shift_y = tan_zen * np.sin(phi) * np.array([alt - zmin, alt - zmax])
    x0, x1 = xmin + shift_x.min(), xmax + shift_x.max()
    y0, y1 = ymin + shift_y.min(), ymax + shift_y.max()

    def samples(lo, hi, step):
        count = int(np.ceil((hi + step - lo) / step))
        return lo + step * np.arange(count)

    xs = samples(x0, x1, x_resolution)
    ys = samples(y0, y1, y_resolution)
    nx, ny = xs.size, ys.size
    px = np.tile(xs, ny)
    py = np.repeat(ys, nx)
    n = px.size
    s = make_sensor_dataset(px, py, np.full(n, alt), np.full(n, mu), np.full(n, phi), stokes,
                            wavelength)
    s["image_shape"] = [nx, ny]
    s["bounding_box"] = np.array([xmin, ymin, zmin, xmax, ymax, zmax])
    s["projection"] = "Orthographic"
    s.update(altitude=alt, x_resolution=x_resolution, y_resolution=y_resolution,
             projection_azimuth=azimuth, projection_zenith=zenith)
    if sub_pixel_ray_args is not None and sub_pixel_ray_args.get("method") is not None:
        ox, oy, w = _subpixel(nx, ny, sub_pixel_ray_args)
        k = w.shape[1]
        _attach_rays(s, px[:, None] + ox * x_resolution / 2.0, py[:, None] + oy * y_resolution / 2.0,
                     np.full((n, k), alt), np.full((n, k), mu), np.full((n, k), phi), w)
    return s


def perspective_projection(wavelength, fov, x_resolution, y_resolution, position_vector,
                           lookat_vector, up_vector, stokes="I", sub_pixel_ray_args=None):
    """A pinhole camera at ``position_vector`` looking at ``lookat_vector``
    with the roll set by ``up_vector``, ``x_resolution`` x ``y_resolution``
    pixels and a field of view of ``fov`` degrees across the longer image
    axis (at3d/sensor.py:367-540).

    Geometry: the camera frame has its third axis w along the viewing
    direction, its first axis u = up x w (normalised) and v = w x u.  The
    image plane lies at focal length f = 1 / tan(fov / 2) with the longer
    axis spanning [-1, 1]; pixel centres sit at the middle of equal cells,
    x fastest.  The ray through image point (a, b) points along
    a/f u + b/f v + w; the sensor fields store the reverse direction (towards
    the camera): mu = -(its z component), phi = its azimuth + pi."""
    nx, ny = int(x_resolution), int(y_resolution)
    if nx != x_resolution or ny != y_resolution or nx < 1 or ny < 1:
        raise ValueError("x_resolution and y_resolution are pixel counts (integers >= 1)")
    # the reference's dtypes are kept (positions and the viewing axis in
    # single precision, the camera's intrinsic matrix inverted in single
    # precision) so that the ray directions are the same to the last bit
    pos = np.asarray(position_vector, np.float32)
    look = np.asarray(lookat_vector, np.float32)
    up = np.asarray(up_vector)
    unit = lambda v: v / np.linalg.norm(v, axis=0)
    wax = unit(look - pos)
    uax = unit(np.cross(up, wax))
    vax = np.cross(wax, uax)
    rot = np.stack((uax, vax, wax), axis=1)
    big = max(nx, ny)
    half_x, half_y = nx / big, ny / big
    cell_x, cell_y = 2.0 * half_x / nx, 2.0 * half_y / ny
    focal = 1.0 / np.tan(np.deg2rad(fov) / 2.0)
    inv_k = np.linalg.inv(np.array([[focal, 0, 0], [0, focal, 0], [0, 0, 1]], dtype=np.float32))
    ax = np.linspace(-half_x + cell_x / 2, half_x - cell_x / 2, nx)
    by = np.linspace(-half_y + cell_y / 2, half_y - cell_y / 2, ny)
    a = np.tile(ax, ny)
    b = np.repeat(by, nx)

    def directions(a, b):
        d = unit(rot @ (inv_k @ np.stack([a, b, np.ones_like(a)])))
        return -d[2].astype(np.float64), (np.arctan2(d[1], d[0]) + np.pi).astype(np.float64)

    mu, phi = directions(a, b)
    n = a.size
    s = make_sensor_dataset(np.full(n, pos[0]), np.full(n, pos[1]), np.full(n, pos[2]), mu, phi,
                            stokes, wavelength)
    s["image_shape"] = [nx, ny]
    s["projection"] = "Perspective"
    s.update(fov_deg=fov, fov_x_deg=float(np.rad2deg(2 * np.arctan(half_x / focal))),
             fov_y_deg=float(np.rad2deg(2 * np.arctan(half_y / focal))), x_resolution=nx,
             y_resolution=ny, position=pos, lookat=look,
             rotation_matrix=np.stack((uax, vax, wax), axis=1).ravel())
    if sub_pixel_ray_args is not None and sub_pixel_ray_args.get("method") is not None:
        ox, oy, w = _subpixel(nx, ny, sub_pixel_ray_args)
        k = w.shape[1]
        ra = (a[:, None] + ox * cell_x / 2.0).ravel()
        rb = (b[:, None] + oy * cell_y / 2.0).ravel()
        rmu, rphi = directions(ra, rb)
        _attach_rays(s, np.full((n, k), pos[0]), np.full((n, k), pos[1]), np.full((n, k), pos[2]),
                     rmu.reshape(n, k), rphi.reshape(n, k), w)
    return s
