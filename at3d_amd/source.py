"""Radiation sources (mirrors at3d/source.py:14-79; solar only)."""
from __future__ import annotations

import numpy as np


def solar(wavelength, solarmu, solar_azimuth, solarflux=1.0, skyrad=0.0):
    """Solar source.  ``solarmu`` points along the propagation of the beam and is
    forced negative; ``solar_azimuth`` is in degrees; ``solarflux`` is the flux
    on a horizontal surface; ``skyrad`` an isotropic downwelling radiance."""
    solarmu = -1 * np.abs(solarmu)
    if not (-1.0 <= solarmu < 0.0):
        raise ValueError("solarmu must be in the range -1.0 <= solarmu < 0.0 not '{}'.".format(solarmu))
    return {
        "name": "solar_source",
        "wavelength": float(wavelength),
        "solarflux": float(solarflux),
        "solarmu": float(solarmu),
        "solaraz": float(np.deg2rad(solar_azimuth)),
        "srctype": "S",
        "units": "R",
        "skyrad": float(skyrad),
    }
