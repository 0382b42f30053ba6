"""Numerical parameters of the solver (mirrors at3d/configuration.py and
default_config.json of the reference).

``get_config()`` returns a plain dict keyed like the reference's xr.Dataset
variables, with the reference's defaults (adaptive cell splitting on at
``split_accuracy`` 0.03).
"""
from __future__ import annotations

import copy

DEFAULTS = {
    "x_boundary_condition": "open",
    "y_boundary_condition": "open",
    "num_mu_bins": 16,
    "num_phi_bins": 32,
    "split_accuracy": 0.03,
    "deltam": True,
    "spherical_harmonics_accuracy": 0.0,
    "acceleration_flag": True,
    "solution_accuracy": 1e-4,
    "max_total_mb": 3000.0,
    "adapt_grid_factor": 5,
    "num_sh_term_factor": 1,
    "cell_to_point_ratio": 1.5,
    "high_order_radiance": False,
    "ip_flag": 0,
    "iterfixsh": 30,
    "tautol": 0.1,
    "angle_set": 2,
    "transcut": 1e-5,
    "transmin": 1.0,
}


def get_config(**overrides) -> dict:
    """Default numerical parameters with keyword overrides (configuration.py:172-203)."""
    cfg = copy.deepcopy(DEFAULTS)
    for k, v in overrides.items():
        if k not in cfg:
            raise KeyError(f"unknown numerical parameter '{k}'")
        cfg[k] = v
    return cfg
