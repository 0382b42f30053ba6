"""Levis approximate gradient of a radiance misfit (mirrors at3d/gradient.py).

``levis_approximation_grad`` is the per-solver, per-sensor entry point
(at3d/gradient.py:164-415); ``LevisApproxGradientUncorrelated`` sums over
solvers and sensors and normalises by the number of measurements
(:431-449); ``levis_loss`` wraps solve + render (forward) and the adjoint
gradient (backward) as a ``torch.autograd.Function``.

Unknowns are given as (scatterer name, variable) pairs with variable
'extinction', 'ssalb' or 'legendre_0_<l>': unit partial derivatives of that
scatterer's optical properties on the property grid with the 'exact' method
(GridToOpticalProperties.calculate_derivatives, at3d/medium.py:1273-1355).
All numerics run in libat3d_hip.so (at3d_levis_gradient).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _max_subgrid_intervals(rte):
    """Sub-step intervals one adjoint ray may record (solver.py:2481-2508 and
    gradient.py:252): the mean cell optical depth over TAUTOL."""
    nx1, ny1, nz = rte.plan_size("nx1"), rte.plan_size("ny1"), rte.plan_size("nz")
    ext = rte.plan_array("total_ext").astype(np.float32).reshape(nx1, ny1, nz)
    cae = (ext[1:, 1:, 1:] + ext[1:, 1:, :-1] + ext[1:, :-1, 1:] + ext[1:, :-1, :-1] +
           ext[:-1, 1:, 1:] + ext[:-1, 1:, :-1] + ext[:-1, :-1, 1:] + ext[:-1, :-1, :-1]) / 8.0
    xg, yg, zg = rte.plan_array("xgrid"), rte.plan_array("ygrid"), rte.plan_array("zgrid")
    vol = (np.diff(zg) * np.diff(xg)[0] * np.diff(yg)[0]) ** (1 / 3)
    tau = vol[np.newaxis, np.newaxis, :] * cae
    return int(np.nanmean(tau) / float(rte.numerical_params["tautol"]))


def _derivative_tables(rte, unknowns):
    """Unit derivative fields and the concatenated derivative Legendre table
    (solver.py:1327-1463)."""
    names = list(rte.medium)
    scs = list(rte.medium.values())
    shape = scs[0].shape
    npg = int(np.prod(shape))
    maxleg = max(s.legcoef.shape[0] for s in scs)
    numder = len(unknowns)
    partder = np.zeros(numder, np.int32)
    dext = np.zeros((numder, npg), np.float32)
    dalb = np.zeros((numder, npg), np.float32)
    diph = np.zeros((numder, npg), np.int32)
    dpw = np.zeros((numder, npg), np.float32)
    tables, offset = [], 0
    for i, (name, var) in enumerate(unknowns):
        if name not in rte.medium:
            raise KeyError(f"unknown scatterer '{name}'")
        isp = names.index(name)
        sc = rte.medium[name]
        partder[i] = isp + 1
        ntab = sc.legcoef.shape[1]
        table = np.zeros((maxleg, ntab))
        if var == "extinction":
            dext[i] = 1.0
        elif var == "ssalb":
            dalb[i] = 1.0
        elif var.startswith("legendre_"):
            parts = var.split("_")
            if len(parts) != 3 or parts[1] != "0":
                raise ValueError(f"variable '{var}': only the I-I component 'legendre_0_<l>' is supported")
            leg = int(parts[2])
            if not 0 <= leg < maxleg:
                raise ValueError(f"Invalid phase component and legendre index for variable name '{var}'")
            table[leg, :] = 1.0
        else:
            raise ValueError(f"variable name is not supported for derivative calculation by this "
                             f"generator.'{var}'")
        diph[i] = np.asarray(sc.table_index, np.int32).ravel() + offset
        dpw[i] = np.asarray(sc.phase_weights, np.float32).ravel()
        tables.append(table)
        offset += ntab
    dleg = np.concatenate(tables, axis=1).T.astype(np.float32)      # [dnumphase][maxleg]
    return partder, dext, dalb, diph, dpw, np.ascontiguousarray(dleg)


def levis_approximation_grad(rte_solver, sensor, unknowns, cost_function="L2",
                             exact_single_scatter=True, indices_for_jacobian=None):
    """Cost and gradient for one solver and one sensor (gradient.py:164-415).

    ``sensor`` holds the ray geometry ('ray_x', 'ray_y', 'ray_z', 'ray_mu',
    'ray_phi'), 'measurement_data' per pixel and optionally 'rays_per_pixel',
    'ray_weight', 'stokes_weights' and 'uncertainties' (the per-pixel inverse
    error variance; 1 by default).  Returns (gradient [numder, npx, npy, npz],
    loss, sensor with the forward radiances in 'I').  With
    ``indices_for_jacobian`` (a tuple of x, y, z index arrays of property grid
    points, as gradient.py:248-252) a fourth value is returned: the Jacobian
    [nstokes=1, numder, npoints, npixels] (float32, the reference's layout)."""
    if cost_function not in ("L2", "LL"):
        raise NotImplementedError("`cost_function` '{}' is not valid.".format(cost_function))
    rte_solver.check_solved()
    rte_solver._ensure_state()
    x = _lib.as_f32(sensor["ray_x"])
    y = _lib.as_f32(sensor["ray_y"])
    z = _lib.as_f32(sensor["ray_z"])
    mu = np.ascontiguousarray(sensor["ray_mu"], dtype=np.float64)
    phi = np.ascontiguousarray(sensor["ray_phi"], dtype=np.float64)
    nrays = x.size
    rpp = np.ascontiguousarray(sensor.get("rays_per_pixel", np.ones(nrays)), np.int32)
    npix = rpp.size
    rw = np.ascontiguousarray(sensor.get("ray_weight", np.ones(nrays)), np.float64)
    sw = np.ascontiguousarray(sensor.get("stokes_weights", np.ones(npix)), np.float64).reshape(-1)
    meas = _lib.as_f32(sensor["measurement_data"]).reshape(-1)
    unc = np.ascontiguousarray(sensor.get("uncertainties", np.ones(npix)), np.float64).reshape(-1)
    if not (meas.size == npix and sw.size == npix and unc.size == npix and rw.size == nrays):
        raise ValueError("pixel arrays must have one entry per pixel and ray weights one per ray")
    partder, dext, dalb, diph, dpw, dleg = _derivative_tables(rte_solver, unknowns)
    numder = partder.size
    dd = _lib.DerivDesc()
    dd.numder = numder
    dd.partder = partder.ctypes.data_as(_lib._i32p)
    dd.dext = dext.ctypes.data_as(_lib._f32p)
    dd.dalb = dalb.ctypes.data_as(_lib._f32p)
    doexact = np.ones(numder, np.int32)
    dd.doexact = doexact.ctypes.data_as(_lib._i32p)
    dd.dnumphase = dleg.shape[0]
    dd.dleg = dleg.ctypes.data_as(_lib._f32p)
    dd.diphasep = diph.ctypes.data_as(_lib._i32p)
    dd.dphasewtp = dpw.ctypes.data_as(_lib._f32p)
    # the derivative table must match the scene's Legendre table length
    if dleg.shape[1] != rte_solver._desc.nlegp + 1:
        raise ValueError("derivative Legendre table length does not match the scene's")
    rte_solver.set_option("max_subgrid_intervals", _max_subgrid_intervals(rte_solver))
    shape = next(iter(rte_solver.medium.values())).shape
    grad = np.zeros((numder,) + tuple(shape), np.float64)
    cost = C.c_double()
    pixels = np.zeros(npix, np.float32)
    err = _lib.errbuf()
    if indices_for_jacobian is None:
        jptr = np.zeros(1, np.int32)
        njac = 0
    else:
        jptr = np.ascontiguousarray(np.ravel_multi_index(tuple(indices_for_jacobian), tuple(shape)),
                                    np.int32).reshape(-1)
        njac = jptr.size
    jac = np.zeros((npix, max(njac, 1), numder), np.float32)
    _lib.check(_lib.lib().at3d_levis_gradient_jacobian(
        rte_solver._state, C.byref(dd), nrays, x.ctypes.data, y.ctypes.data, z.ctypes.data,
        mu.ctypes.data, phi.ctypes.data, npix, rpp.ctypes.data, rw.ctypes.data, sw.ctypes.data,
        meas.ctypes.data, unc.ctypes.data, cost_function.encode(), int(bool(exact_single_scatter)),
        njac, jptr.ctypes.data, jac.ctypes.data, grad.ctypes.data, C.byref(cost),
        pixels.ctypes.data, err, len(err)), err)
    out = dict(sensor)
    out["I"] = pixels
    if indices_for_jacobian is None:
        return grad, float(cost.value), out
    # [npix][njac][numder] -> (nstokes, numder, njac, npix)
    return grad, float(cost.value), out, np.transpose(jac, (2, 1, 0))[np.newaxis].copy()


class LevisApproxGradientUncorrelated:
    """Sum of the per-sensor Levis gradients over solvers, normalised by the
    number of measurements (gradient.py:431-449).

    ``measurements`` maps a solver key to a list of sensors carrying
    'measurement_data'; ``solvers`` maps the same keys to RTE objects;
    ``unknowns`` is a list of (scatterer name, variable) pairs;
    ``gradient_kwargs`` holds 'cost_function' and optionally
    'exact_single_scatter'; ``solve_kwargs`` is passed to RTE.solve."""

    def __init__(self, measurements, solvers, unknowns, gradient_kwargs, solve_kwargs=None):
        if "cost_function" not in gradient_kwargs:
            raise ValueError("'cost_function' must be specified in `gradient_kwargs`. Supported values "
                             "are 'L2' or 'LL'.")
        self.measurements = measurements
        self.solvers = solvers
        self.unknowns = list(unknowns)
        self.gradient_kwargs = dict(gradient_kwargs)
        self.solve_kwargs = dict(solve_kwargs or {"maxiter": 100, "verbose": False})

    @property
    def nmeasurements(self):
        return sum(int(np.size(s["measurement_data"])) for sens in self.measurements.values()
                   for s in sens)

    def __call__(self):
        loss, grad = 0.0, None
        for key, rte in self.solvers.items():
            rte.solve(**self.solve_kwargs)
            for sensor in self.measurements[key]:
                g, c, _ = levis_approximation_grad(
                    rte, sensor, self.unknowns,
                    cost_function=self.gradient_kwargs["cost_function"],
                    exact_single_scatter=self.gradient_kwargs.get("exact_single_scatter", True))
                loss += c
                grad = g if grad is None else grad + g
        n = self.nmeasurements
        # (x, y, z, derivative_index) as make_gradient_dataset (gradient.py:451-504)
        return loss / n, np.moveaxis(grad / n, 0, -1)


def levis_loss(extinction, ssalb, make_rte, sensors, scatterer, cost_function="L2",
               exact_single_scatter=True, solve_kwargs=None):
    """L2/LL misfit of rendered radiances as a differentiable torch function of
    one scatterer's extinction and single-scattering albedo (property-grid
    tensors).  ``make_rte(extinction, ssalb)`` builds an RTE from numpy arrays;
    ``sensors`` carry 'measurement_data'.  Forward: solve + render; backward:
    the Levis approximate gradient, normalised by the number of measurements
    as LevisApproxGradientUncorrelated."""
    import torch

    class _Levis(torch.autograd.Function):
        @staticmethod
        def forward(ctx, ext, alb):
            rte = make_rte(ext.detach().cpu().numpy(), alb.detach().cpu().numpy())
            rte.solve(**(solve_kwargs or {"maxiter": 100, "verbose": False}))
            loss, grad, n = 0.0, None, 0
            for sensor in sensors:
                g, c, _ = levis_approximation_grad(rte, sensor, [(scatterer, "extinction"),
                                                                 (scatterer, "ssalb")],
                                                   cost_function=cost_function,
                                                   exact_single_scatter=exact_single_scatter)
                loss += c
                grad = g if grad is None else grad + g
                n += int(np.size(sensor["measurement_data"]))
            ctx.save_for_backward(torch.as_tensor(grad[0] / n, dtype=ext.dtype, device=ext.device),
                                  torch.as_tensor(grad[1] / n, dtype=alb.dtype, device=alb.device))
            rte.release()
            return torch.as_tensor(loss / n, dtype=ext.dtype, device=ext.device)

        @staticmethod
        def backward(ctx, gout):
            ge, ga = ctx.saved_tensors
            return gout * ge, gout * ga

    return _Levis.apply(extinction, ssalb)
