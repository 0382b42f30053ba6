"""Surfaces (mirrors at3d/surface.py:28-105; fixed Lambertian only)."""
from __future__ import annotations


def lambertian(albedo, ground_temperature=298.15):
    """Fixed Lambertian surface with reflectance ``albedo`` in [0, 1]."""
    albedo = float(albedo)
    if not 0.0 <= albedo <= 1.0:
        raise ValueError("surface albedo should be in [0, 1] not '{}'".format(albedo))
    return {"name": "fixed_lambertian_surface", "sfctype": "FL", "gndalbedo": albedo,
            "gndtemp": float(ground_temperature)}
