"""Gridded optical properties of one scattering species.

Mirrors the fields of the reference's scatterer xr.Dataset
(at3d/solver.py:1689-1760 checks; medium.py): ``extinction`` and ``ssalb`` on
an (x, y, z) grid, Legendre tables ``legcoef`` (legendre_index, table_index)
including the (2l+1) factor, or (stokes_index, legendre_index, table_index)
with the six phase-matrix elements a1, a2, a3, a4, b1, b2 for polarized
solves, 1-based ``table_index`` and ``phase_weights``.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Scatterer:
    extinction: np.ndarray                 # (nx, ny, nz) 1/km
    ssalb: np.ndarray                      # (nx, ny, nz)
    legcoef: np.ndarray                    # (nleg+1, ntable) I-I, or (6, nleg+1, ntable)
    delx: float
    dely: float
    z: np.ndarray                          # (nz,) km
    table_index: np.ndarray | None = None  # (nx, ny, nz) or (num_micro, nx, ny, nz), 1-based
    phase_weights: np.ndarray | None = None  # the same shape: weights of the tables mixed at a point

    def __post_init__(self):
        self.extinction = np.asarray(self.extinction, np.float32)
        self.ssalb = np.asarray(self.ssalb, np.float32)
        leg = np.asarray(self.legcoef, np.float64)
        if leg.ndim == 1:
            leg = leg[:, None]
        if leg.ndim == 3 and leg.shape[0] != 6:
            raise ValueError("a polarized Legendre table has six elements (stokes_index)")
        self.legcoef = leg
        self.z = np.asarray(self.z, np.float32)
        shp = self.extinction.shape
        if self.ssalb.shape != shp or len(shp) != 3 or self.z.shape != (shp[2],):
            raise ValueError("extinction, ssalb and z must describe one (x, y, z) grid")
        if self.table_index is None:
            self.table_index = np.ones(shp, np.int32)
        if self.phase_weights is None:
            self.phase_weights = np.ones(np.shape(self.table_index), np.float32)
        self.table_index = np.asarray(self.table_index, np.int32)
        self.phase_weights = np.asarray(self.phase_weights, np.float32)
        if self.table_index.shape != self.phase_weights.shape or \
                self.table_index.shape[-3:] != shp or self.table_index.ndim not in (3, 4):
            raise ValueError("table_index and phase_weights must be (x, y, z) or (num_micro, x, y, z) arrays")
        if np.any(self.extinction < 0) or np.any(self.ssalb < 0) or np.any(self.ssalb > 1):
            raise ValueError("extinction must be >= 0 and ssalb in [0, 1]")

    @property
    def shape(self):
        return self.extinction.shape

    @property
    def num_micro(self):
        """Phase tables mixed at each grid point (medium.py's num_micro)."""
        return 1 if self.table_index.ndim == 3 else self.table_index.shape[0]
