"""Levis approximate gradient of a radiance misfit (mirrors at3d/gradient.py).

``levis_approximation_grad`` is the per-solver, per-sensor entry point
(at3d/gradient.py:164-415); ``LevisApproxGradient`` and
``LevisApproxGradientUncorrelated`` keep the reference's constructor and
return values (:21-162, :417-449), with ``make_gradient_dataset`` and
``make_jacobian_dataset`` (:451-579) giving dicts of arrays in place of
xarray datasets; ``levis_loss`` wraps solve + render (forward) and the
adjoint gradient (backward) as a ``torch.autograd.Function``.

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


class NullUncertainty:
    """Unweighted least squares (at3d/uncertainties.py:109-140): every pixel's
    1 x 1 inverse error covariance is ``scaling_factor``; adds no noise."""

    def __init__(self, cost_function, scaling_factor=1.0):
        self.cost_function = cost_function
        self._scaling_factor = float(scaling_factor)

    def calculate_uncertainties(self, sensor):
        npix = np.size(sensor["measurement_data"])
        sensor["uncertainties"] = self._scaling_factor * np.ones(npix)
        sensor["num_uncertainty"] = 1

    def add_noise(self, sensor):
        raise ValueError("{} cannot be used to generate measurement noise. Please assign "
                         "another uncertainty model.".format(type(self)))


def _variables(unknown):
    return list(getattr(unknown, "variables", unknown))


def _unknown_list(unknown_scatterers):
    return [(name, var) for name, u in unknown_scatterers.items() for var in _variables(u)]


class LevisApproxGradient:
    """Prepares and evaluates the Levis approximate gradient over solvers and
    sensors (at3d/gradient.py:21-162, 164-415).

    ``measurements`` and ``forward_sensors`` map an instrument name to
    {'sensor_list': [sensor, ...], 'uncertainty_model': model or None}
    (the reference's SensorsDict): a measured sensor carries
    'measurement_data' (and 'wavelength', 'ray_*', optionally
    'rays_per_pixel', 'ray_weight', 'stokes_weights'); its forward sensor at
    the same position receives the modelled pixels in 'I'.  ``solvers`` maps a
    wavelength to an RTE (the reference's SolversDict); each sensor is
    evaluated with the solver of its wavelength.  ``unknown_scatterers`` maps a
    scatterer name to its unknown variables ('extinction', 'ssalb',
    'legendre_0_<l>'; a list or an object with ``.variables``).
    ``parallel_solve_kwargs`` go to RTE.solve ('n_jobs' and 'mpi_comm' are
    accepted and unused: one GPU evaluates every ray);
    ``gradient_kwargs`` need 'cost_function' ('L2' or 'LL') and may hold
    'exact_single_scatter' and 'indices_for_jacobian'; ``uncertainty_kwargs``
    'add_noise' (default False, with a warning)."""

    def __init__(self, measurements, solvers, forward_sensors, unknown_scatterers,
                 parallel_solve_kwargs, gradient_kwargs, uncertainty_kwargs):
        import warnings
        if "add_noise" not in uncertainty_kwargs:
            uncertainty_kwargs["add_noise"] = False
            warnings.warn("'add_noise' flag was unspecified. It has been set to False.")
        elif not isinstance(uncertainty_kwargs["add_noise"], bool):
            raise TypeError("uncertainty_kwargs['add_noise'] should be of boolean type.")
        if "cost_function" not in gradient_kwargs:
            raise ValueError("'cost_function' must be specified in `gradient_kwargs`. Supported values "
                             "are 'L2' or 'LL'.")
        if not isinstance(gradient_kwargs["cost_function"], str):
            raise TypeError("gradient_kwargs['cost_function'] should be of string type.")
        self.measurements = measurements
        self.solvers = solvers
        self.forward_sensors = forward_sensors
        self.unknown_scatterers = unknown_scatterers
        self.parallel_solve_kwargs = parallel_solve_kwargs
        self.gradient_kwargs = gradient_kwargs
        self.uncertainty_kwargs = uncertainty_kwargs
        for name, instrument in self.measurements.items():
            if instrument.get("uncertainty_model") is None:
                warnings.warn("No uncertainty model supplied for instrument '{}'. Using "
                              "NullUncertainty which is equivalent to unweighted least squares."
                              .format(name))
                instrument["uncertainty_model"] = NullUncertainty(gradient_kwargs["cost_function"])
            model = instrument["uncertainty_model"]
            if model.cost_function != gradient_kwargs["cost_function"]:
                raise ValueError("Uncertainty model's assumed cost_function '{}' is inconsistent with "
                                 "the one being used '{}'".format(model.cost_function,
                                                                  gradient_kwargs["cost_function"]))
            for sensor in instrument["sensor_list"]:
                model.calculate_uncertainties(sensor)
                if uncertainty_kwargs["add_noise"]:
                    model.add_noise(sensor)

    @property
    def nmeasurements(self):
        return sum(int(np.size(s["measurement_data"])) for inst in self.measurements.values()
                   for s in inst["sensor_list"])

    def _solver_for(self, sensor):
        wl = float(sensor["wavelength"])
        for key, rte in self.solvers.items():
            if np.isclose(float(key), wl):
                return rte
        raise KeyError("no solver for the sensor's wavelength {}".format(wl))

    def _prep_gradient(self):
        """Solve every solver, then per measured sensor the cost, gradient
        and forward pixels (and the Jacobian when asked).  Returns
        (losses, gradients [numder, npx, npy, npz], jacobians)."""
        solve_kwargs = {k: v for k, v in self.parallel_solve_kwargs.items() if k not in ("n_jobs", "mpi_comm")}
        solve_kwargs.setdefault("maxiter", 100)
        solve_kwargs.setdefault("verbose", False)
        for rte in self.solvers.values():
            if not rte.check_solved(verbose=False):
                rte.solve(**solve_kwargs)
        gk = {k: v for k, v in self.gradient_kwargs.items()}
        losses, grads, jacobians = [], [], []
        for name, instrument in self.measurements.items():
            forward = self.forward_sensors[name]["sensor_list"]
            for i, meas in enumerate(instrument["sensor_list"]):
                rte = self._solver_for(meas)
                sensor = dict(forward[i])
                for k in ("measurement_data", "uncertainties", "stokes_weights"):
                    if k in meas:
                        sensor[k] = meas[k]
                g, c, rays, jac = self.levis_approximation_grad(
                    rte, sensor, cost_function=gk["cost_function"],
                    indices_for_jacobian=gk.get("indices_for_jacobian"),
                    exact_single_scatter=gk.get("exact_single_scatter", True))
                forward[i]["I"] = rays["I"]
                losses.append(c)
                grads.append(g)
                if jac is not None:
                    jacobians.append((float(meas["wavelength"]), jac))
        return losses, grads, jacobians

    def levis_approximation_grad(self, rte_solver, sensor, cost_function="L2",
                                 indices_for_jacobian=None, exact_single_scatter=True):
        """(gradient, loss, integrated_rays, jacobian) for one solver and one
        sensor (gradient.py:164-415); jacobian is None without
        ``indices_for_jacobian``."""
        unknowns = [u for u in _unknown_list(self.unknown_scatterers) if u[0] in rte_solver.medium]
        out = levis_approximation_grad(rte_solver, sensor, unknowns, cost_function=cost_function,
                                       exact_single_scatter=exact_single_scatter,
                                       indices_for_jacobian=indices_for_jacobian)
        if indices_for_jacobian is None:
            g, c, rays = out
            return g, c, rays, None
        g, c, rays, jac = out
        return g, c, rays, jac

    def __call__(self):
        return self._prep_gradient(), None, None


class LevisApproxGradientUncorrelated(LevisApproxGradient):
    """Different sensors' errors uncorrelated: the loss and gradient are the
    sums over sensors divided by the number of measurements
    (gradient.py:417-449).  Returns (loss, gradient_dataset,
    jacobian_dataset)."""

    def __call__(self):
        losses, grads, jacobians = self._prep_gradient()
        n = self.nmeasurements
        loss = float(np.sum(losses)) / n
        gradient = np.sum(grads, axis=0) / n
        gradient_dataset = make_gradient_dataset(gradient, self.unknown_scatterers, self.solvers)
        jacobian_dataset = None
        if jacobians:
            jacobian_dataset = make_jacobian_dataset(jacobians, self.unknown_scatterers,
                                                     self.gradient_kwargs["indices_for_jacobian"],
                                                     self.solvers)
        return loss, gradient_dataset, jacobian_dataset


def _grid_coords(solvers):
    rte = next(iter(solvers.values()))
    sc = next(iter(rte.medium.values()))
    nx, ny, nz = sc.shape
    return (np.arange(nx) * sc.delx, np.arange(ny) * sc.dely, np.asarray(sc.z))


def make_gradient_dataset(gradient, unknown_scatterers, solvers):
    """The gradient on the property grid with its coordinates and the
    (scatterer, variable) of each derivative index (gradient.py:451-504):
    {'gradient': (x, y, z, derivative_index), 'x', 'y', 'z',
    'derivative_index'}.  ``gradient`` is [numder, npx, npy, npz]."""
    x, y, z = _grid_coords(solvers)
    g = np.moveaxis(np.asarray(gradient), 0, -1)
    return {"gradient": g.reshape((x.size, y.size, z.size, -1)), "x": x, "y": y, "z": z,
            "derivative_index": _unknown_list(unknown_scatterers)}


def make_jacobian_dataset(jacobian_list, unknown_scatterers, indices_for_jacobian, solvers):
    """Frechet derivatives per solver wavelength (gradient.py:506-579):
    {'jacobian_<wavelength>': (nstokes, derivative_index, grid_index,
    npixels of that wavelength's sensors), 'grid_index': the (x, y, z)
    coordinates of the requested points, 'derivative_index'}.
    ``jacobian_list`` holds (wavelength, jacobian) per evaluated sensor."""
    x, y, z = _grid_coords(solvers)
    ix = tuple(np.asarray(a) for a in indices_for_jacobian)
    out = {"grid_index": list(zip(x[ix[0]], y[ix[1]], z[ix[2]])),
           "derivative_index": _unknown_list(unknown_scatterers)}
    for key in solvers:
        parts = [j for wl, j in jacobian_list if np.isclose(wl, float(key))]
        if parts:
            out["jacobian_{:1.3f}".format(float(key))] = np.concatenate(parts, axis=-1)
    return out


def _pixels_and_cost(rte, sensor, cost_function):
    """Forward pixels (ray- and Stokes-weighted sums of rendered rays) and the
    cost of one sensor, as COMPUTE_ADJOINT_WEIGHTS forms them
    (shdomsub4.f:3969-4034): L2 0.5 U (R - M)^2, LL 0.5 U (log R - log M)^2."""
    rays = rte.integrate_to_sensor({k: sensor[k] for k in ("ray_x", "ray_y", "ray_z", "ray_mu",
                                                          "ray_phi")})["I"]
    nrays = rays.size
    rpp = np.asarray(sensor.get("rays_per_pixel", np.ones(nrays)), np.int64)
    rw = np.asarray(sensor.get("ray_weight", np.ones(nrays)), np.float64)
    sw = np.asarray(sensor.get("stokes_weights", np.ones(rpp.size)), np.float64).reshape(-1)
    unc = np.asarray(sensor.get("uncertainties", np.ones(rpp.size)), np.float64).reshape(-1)
    meas = np.asarray(sensor["measurement_data"], np.float32).reshape(-1).astype(np.float64)
    if np.all(rpp == 1):
        pix = (rays.astype(np.float64) * rw * sw).astype(np.float32)
    else:
        # the pixel's running sum rounded to FP32 after each ray, as there
        pix = np.zeros(rpp.size, np.float32)
        k = 0
        for ip, n in enumerate(rpp):
            so = np.float32(0.0)
            for _ in range(n):
                so = np.float32(float(so) + float(rays[k]) * rw[k] * sw[ip])
                k += 1
            pix[ip] = so
    r = pix.astype(np.float64)
    e = (r - meas) if cost_function == "L2" else (np.log(r) - np.log(meas))
    return pix, float(np.sum(0.5 * unc * e * e))


def levis_loss(extinction, ssalb, make_rte, sensors, scatterer, cost_function="L2",
               exact_single_scatter=True, solve_kwargs=None):
    """L2/LL misfit of rendered radiances as a differentiable torch function of
    one scatterer's extinction and single-scattering albedo (property-grid
    tensors).  ``make_rte(extinction, ssalb)`` builds an RTE from numpy arrays;
    ``sensors`` carry 'measurement_data'.  Forward: solve + render and the
    cost, normalised by the number of measurements as
    LevisApproxGradientUncorrelated; backward: the Levis approximate adjoint
    (at3d_levis_gradient) on the solution the forward kept, times the
    incoming gradient.  The solution is released after backward (or when the
    graph is freed)."""
    import torch

    if cost_function not in ("L2", "LL"):
        raise NotImplementedError("`cost_function` '{}' is not valid.".format(cost_function))

    class _Levis(torch.autograd.Function):
        @staticmethod
        def forward(ctx, ext, alb):
            rte = make_rte(ext.detach().cpu().numpy(), alb.detach().cpu().numpy())
            rte.solve(**(solve_kwargs or {"maxiter": 100, "verbose": False}))
            loss, n = 0.0, 0
            for sensor in sensors:
                _, c = _pixels_and_cost(rte, sensor, cost_function)
                loss += c
                n += int(np.size(sensor["measurement_data"]))
            ctx.rte, ctx.n = rte, n
            return torch.as_tensor(loss / n, dtype=ext.dtype, device=ext.device)

        @staticmethod
        def backward(ctx, gout):
            rte, n = ctx.rte, ctx.n
            grad = None
            for sensor in sensors:
                g, _, _ = levis_approximation_grad(rte, sensor, [(scatterer, "extinction"),
                                                                 (scatterer, "ssalb")],
                                                   cost_function=cost_function,
                                                   exact_single_scatter=exact_single_scatter)
                grad = g if grad is None else grad + g
            rte.release()
            ctx.rte = None
            ge = torch.as_tensor(grad[0] / n, dtype=gout.dtype, device=gout.device)
            ga = torch.as_tensor(grad[1] / n, dtype=gout.dtype, device=gout.device)
            return gout * ge, gout * ga

    return _Levis.apply(extinction, ssalb)
