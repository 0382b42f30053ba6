"""Collections of sensors and solvers (at3d/containers.py:39-133, :674-768).

``SolversDict`` maps a key (the monochromatic wavelength) to an RTE and solves
the ones that need it, on host threads when ``n_jobs > 1`` (each RTE owns a
device state and a HIP stream, so the solves overlap on the GPU).
``SensorsDict`` holds sensors (the dicts of at3d_amd.sensor) by instrument
and renders them with the solvers of their wavelengths: the Stokes
components a sensor asks for, per ray, then averaged to its pixels
(AVERAGE_SUBPIXEL_RAYS) as get_measurements does.  Plain dicts stand in for
the reference's xr.Dataset sensors; MPI is not used.
"""
from __future__ import annotations

from collections import OrderedDict
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import sensor as _sensor
from .solver import RTE

_RAY_KEYS = ("ray_x", "ray_y", "ray_z", "ray_mu", "ray_phi")


class SolversDict(OrderedDict):
    """RTE objects by key (at3d/containers.py:674)."""

    def add_solver(self, key, solver):
        if not isinstance(solver, RTE):
            raise TypeError("solver should be of type '{}'".format(RTE))
        self[key] = solver

    def to_solve(self, overwrite_solver):
        """(keys, solvers) that need a solve: all with ``overwrite_solver``,
        else those not converged (containers.py:746-767)."""
        keys = [k for k, s in self.items() if overwrite_solver or not s.check_solved(verbose=False)]
        return keys, [self[k] for k in keys]

    def solve(self, n_jobs=1, mpi_comm=None, overwrite_solver=False, maxiter=100, verbose=True,
              init_solution=True, setup_grid=True):
        """Solves every solver that needs it (containers.py:699-744); with
        n_jobs > 1 on that many host threads."""
        if mpi_comm is not None:
            raise NotImplementedError("MPI distribution is not supported: use n_jobs")
        _, to_solve = self.to_solve(overwrite_solver)

        def run(s):
            s.solve(maxiter=maxiter, init_solution=init_solution, verbose=verbose)
        if n_jobs == 1 or len(to_solve) <= 1:
            for s in to_solve:
                run(s)
        else:
            with ThreadPoolExecutor(min(int(n_jobs), len(to_solve))) as ex:
                list(ex.map(run, to_solve))


class SensorsDict(OrderedDict):
    """Sensors by instrument (at3d/containers.py:39)."""

    def add_sensor(self, instrument, sensor):
        for k in _RAY_KEYS + ("wavelength", "stokes", "pixel_index", "ray_weight"):
            if k not in sensor:
                raise KeyError("sensor is missing '{}'".format(k))
        if instrument not in self:
            self[instrument] = {"sensor_list": [], "uncertainty_model": None}
        self[instrument]["sensor_list"].append(sensor)

    def add_uncertainty_model(self, instrument, uncertainty_model):
        self[instrument]["uncertainty_model"] = uncertainty_model

    def get_unique_solvers(self):
        """The wavelengths the sensors need (containers.py:355)."""
        return sorted({float(s["wavelength"]) for inst in self.values() for s in inst["sensor_list"]})

    def get_measurements(self, solvers, n_jobs=1, mpi_comm=None, maxiter=100, verbose=True,
                         init_solution=True, setup_grid=True, destructive=False, overwrite_solver=False):
        """Solves what needs solving, renders every sensor's rays with the
        solver of its wavelength and stores each requested Stokes component
        per pixel in the sensor (the ray values under 'ray_<S>')
        (containers.py:133-233)."""
        missing = [w for w in self.get_unique_solvers() if w not in solvers]
        if missing:
            raise KeyError("no solver for the wavelengths {}".format(missing))
        solvers.solve(n_jobs=n_jobs, mpi_comm=mpi_comm, overwrite_solver=overwrite_solver,
                      maxiter=maxiter, verbose=verbose, init_solution=init_solution)
        for inst in self.values():
            for s in inst["sensor_list"]:
                rte = solvers[float(s["wavelength"])]
                rays = {k: s[k] for k in _RAY_KEYS}
                rays["stokes"] = s["stokes"]
                out = rte.integrate_to_sensor(rays)
                for comp in s["stokes"]:
                    s["ray_" + comp] = np.asarray(out[comp])
                    s[comp] = _sensor.average_subpixel_rays(s, out[comp]).astype(np.float32)
        return self
