"""Sensor geometry (mirrors at3d/sensor.py: make_sensor_dataset :31,
orthographic_projection :188) as plain dicts of numpy arrays.

A sensor holds per-pixel 'cam_*' arrays and per-ray 'ray_*' arrays (one ray
per pixel here: sub-pixel ray averaging is not implemented), with 'ray_mu'
the cosine zenith and 'ray_phi' the azimuth (radians) of the direction TOWARDS
the sensor.
"""
from __future__ import annotations

import itertools

import numpy as np


def make_sensor_dataset(x, y, z, mu, phi, stokes, wavelength):
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
    stokes = list(np.atleast_1d(stokes))
    if stokes != ["I"]:
        raise ValueError("only the 'I' Stokes component is supported (num_stokes=1)")
    s = {"wavelength": float(wavelength), "stokes": stokes,
         "cam_x": x.astype(np.float64), "cam_y": y.astype(np.float64),
         "cam_z": z.astype(np.float64), "cam_mu": mu.astype(np.float64),
         "cam_phi": phi.astype(np.float64)}
    for k in ("x", "y", "z", "mu", "phi"):
        s["ray_" + k] = s["cam_" + k].copy()
    s["pixel_index"] = np.arange(x.size)
    s["ray_weight"] = np.ones(x.size)
    return s


def orthographic_projection(wavelength, bounding_box, x_resolution, y_resolution, azimuth, zenith,
                            altitude="TOA", stokes="I"):
    """Parallel rays viewing ``bounding_box`` = (xmin, ymin, zmin, xmax, ymax, zmax)
    at azimuth/zenith (degrees, direction of photons), sampled every
    x_resolution / y_resolution km on the projected footprint, padded to cover
    the domain edge (at3d/sensor.py:188-347 without sub-pixel rays)."""
    mu = np.cos(np.deg2rad(zenith))
    phi = np.deg2rad(azimuth)
    xmin, ymin, zmin, xmax, ymax, zmax = bounding_box
    altitude = zmax if altitude == "TOA" else altitude
    alpha = np.sqrt(1 - mu ** 2) * np.cos(phi) / mu
    beta = np.sqrt(1 - mu ** 2) * np.sin(phi) / mu
    projection_matrix = np.array([[1, 0, -alpha, alpha * altitude],
                                  [0, 1, -beta, beta * altitude],
                                  [0, 0, 0, altitude]])
    pts = np.array(list(itertools.product([xmin, xmax], [ymin, ymax], [zmin, zmax]))).T
    proj = projection_matrix @ np.pad(pts, ((0, 1), (0, 0)), constant_values=1)
    x_s, y_s = proj[:2, :].min(axis=1)
    x_e, y_e = proj[:2, :].max(axis=1)
    x = np.arange(x_s, x_e + x_resolution, x_resolution)
    y = np.arange(y_s, y_e + y_resolution, y_resolution)
    z = altitude
    image_shape = [x.size, y.size]
    x, y, z, mu, phi = np.meshgrid(x, y, z, mu, phi)
    s = make_sensor_dataset(x.ravel(), y.ravel(), z.ravel(), mu.ravel(), phi.ravel(), stokes,
                            wavelength)
    s["image_shape"] = image_shape
    s["projection"] = "Orthographic"
    return s
