"""The SHDOM radiative-transfer solver object on one MI355X.

``RTE`` mirrors the reference's ``at3d.solver.RTE`` (at3d/solver.py:143-2951)
for the scalar (NSTOKES=1) solar path on a fixed base grid: the constructor
prepares the grid, optics and discrete ordinates, ``solve`` runs the
source-function iteration on the GPU, ``integrate_to_sensor`` renders rays.
All numerics run in libat3d_hip.so behind the C ABI of include/at3d_hip.h;
this module only marshals arrays.
"""
from __future__ import annotations

import ctypes as C
import warnings
from collections import OrderedDict

import numpy as np

from . import _lib
from .configuration import get_config
from .medium import Scatterer


def device_trim() -> int:
    """Return the device blocks kept from destroyed solver states to the
    runtime; gives the bytes they held."""
    return int(_lib.lib().at3d_device_trim())


class RTE:
    """Radiative transfer solver (at3d/solver.py:143)."""

    def __init__(self, numerical_params, medium, source, surface, num_stokes=1, name=None,
                 atmosphere=None, device_setup=True, interpmethod="ON"):
        if num_stokes not in (1, 3):
            raise ValueError("num_stokes must be 1 or 3 (NSTOKES = 4 is not supported by this engine)")
        if atmosphere is not None:
            raise ValueError("atmosphere (thermal / gas absorption) is not supported")
        self._nstokes = int(num_stokes)
        # INTERPMETHOD of the property transfer (solver.py:156): 'ON' mixes the
        # eight property points' tables, 'OO' takes the one of the point with
        # the most scattering (the reference's property-file verification)
        if interpmethod not in ("ON", "OO"):
            raise ValueError("interpmethod must be 'ON' or 'OO'")
        self._interpmethod = interpmethod
        self._phasemax = 0.999
        # RENDER's CORRECTINTERPOLATE (solver.py:170)
        self._correctinterpolate = True
        self.numerical_params = dict(get_config()) if numerical_params is None else dict(numerical_params)
        self.source = dict(source)
        self.surface = dict(surface)
        if self.source.get("srctype", "S") != "S":
            raise ValueError("only solar sources are supported")
        if self.surface.get("sfctype", "FL") != "FL":
            raise ValueError("only fixed Lambertian surfaces are supported")
        self.medium = self._setup_medium(medium)
        self.wavelength = float(self.source["wavelength"])
        self._name = "Radiance {:1.3f} micron".format(self.wavelength) if name is None else name
        self._solcrit = None
        self._iters = 0
        self._state = None
        self._plan = None
        self._restore_data = None   # a loaded solution (load_solution)
        self._keep = []          # arrays referenced by the C descriptor
        # the per-point medium (property transfer, delta-M, constants,
        # two-stream start) made on the device with the solver state rather
        # than on the host with the plan; the same values either way
        self._device_setup = bool(device_setup)
        self._build_plan()

    # -- setup -----------------------------------------------------------------
    @staticmethod
    def _setup_medium(medium):
        if isinstance(medium, Scatterer):
            md = OrderedDict(scatterer_000=medium)
        elif isinstance(medium, dict):
            md = OrderedDict(medium)
        else:
            md = OrderedDict(("scatterer_{:03d}".format(i), v) for i, v in enumerate(medium))
        first = next(iter(md.values()))
        for k, v in md.items():
            if not isinstance(v, Scatterer):
                raise TypeError(f"medium['{k}'] must be an at3d_amd.medium.Scatterer")
            if v.shape != first.shape or not np.allclose(v.z, first.z) or \
                    v.delx != first.delx or v.dely != first.dely:
                raise ValueError("all scatterers must be on the same grid")
        return md

    def _build_plan(self):
        npar = self.numerical_params
        scs = list(self.medium.values())
        npx, npy, npz = scs[0].shape
        # concatenated Legendre table (solver.py:2340-2411): the I-I element
        # for a scalar solve, all six elements (a scalar table is a1 alone)
        # for a polarized one
        nstleg = 1 if self._nstokes == 1 else 6

        def elements(leg):
            if leg.ndim == 2:
                leg = np.concatenate([leg[None], np.zeros((5,) + leg.shape)], axis=0)
            return leg[:nstleg]
        legs = [elements(s.legcoef) for s in scs]
        maxleg = max(t.shape[1] for t in legs)
        tables, offset = [], 0
        iphasep = []
        # phase tables per grid point (num_micro): every species padded to the
        # largest count with zero weights (solver.py:2340-2375)
        nmicro = max(s.num_micro for s in scs)

        def micro(a, fill):
            a = np.asarray(a)
            if a.ndim == 3:
                a = a[None]
            if a.shape[0] < nmicro:
                a = np.concatenate([a, np.full((nmicro - a.shape[0],) + a.shape[1:], fill, a.dtype)])
            return np.moveaxis(a, 0, -1)   # (x, y, z, num_micro)
        # each species' pointers are offset by the largest pointer of the
        # species before it (solver.py:2353-2356); pointers of 0 become 1
        pws = []
        for s, t in zip(scs, legs):
            t = np.pad(t, ((0, 0), (0, maxleg - t.shape[1]), (0, 0)))
            tables.append(t)
            ti = micro(np.asarray(s.table_index, np.int32), 1) + offset
            iphasep.append(ti)
            pws.append(micro(np.asarray(s.phase_weights, np.float32), 0.0))
            offset = max(offset, int(ti.max()))
        table3 = np.concatenate(tables, axis=2)           # (nstleg, nleg+1, numphase)
        table = table3[0]                                 # (nleg+1, numphase)
        legenp = _lib.as_f32(np.transpose(table3, (2, 1, 0)))   # [numphase][nlegp+1][nstleg]
        ext = _lib.as_f32(np.stack([s.extinction for s in scs]))
        alb = _lib.as_f32(np.stack([s.ssalb for s in scs]))
        iph = np.stack(iphasep)
        iph[iph == 0] = 1
        iph = _lib.as_i32(iph)
        pw = _lib.as_f32(np.stack(pws))
        zl = _lib.as_f32(scs[0].z)
        bcx, bcy = npar["x_boundary_condition"], npar["y_boundary_condition"]
        if bcx != bcy:
            raise ValueError("x and y boundary conditions must match (both 'open' or 'periodic')")
        d = _lib.SceneDesc()
        d.npx, d.npy, d.npz = npx, npy, npz
        d.delx, d.dely = float(scs[0].delx), float(scs[0].dely)
        d.zlevels = zl.ctypes.data_as(_lib._f32p)
        d.npart = len(scs)
        d.extinctp = ext.ctypes.data_as(_lib._f32p)
        d.albedop = alb.ctypes.data_as(_lib._f32p)
        d.iphasep = iph.ctypes.data_as(_lib._i32p)
        d.phasewtp = pw.ctypes.data_as(_lib._f32p)
        d.numphase = table.shape[1]
        d.nlegp = table.shape[0] - 1
        d.legenp = legenp.ctypes.data_as(_lib._f32p)
        src, sfc = self.source, self.surface
        d.solarflux, d.solarmu, d.solaraz = src["solarflux"], src["solarmu"], src["solaraz"]
        d.skyrad, d.gndalbedo, d.wavelength = src.get("skyrad", 0.0), sfc["gndalbedo"], self.wavelength
        d.nmu, d.nphi = int(npar["num_mu_bins"]), int(npar["num_phi_bins"])
        d.open_bc = 1 if bcx == "open" else 0
        d.ip_flag = int(npar["ip_flag"])
        d.deltam = int(bool(npar["deltam"]))
        d.accelflag = int(bool(npar["acceleration_flag"]))
        d.high_order_radiance = int(bool(npar["high_order_radiance"]))
        d.iterfixsh = int(npar["iterfixsh"])
        d.angle_set = int(npar["angle_set"])
        d.shacc = float(npar["spherical_harmonics_accuracy"])
        d.solacc = float(npar["solution_accuracy"])
        d.splitacc = float(npar["split_accuracy"])
        d.transmin = float(npar["transmin"])
        d.tautol = float(npar["tautol"])
        d.transcut = float(npar["transcut"])
        d.maxig, d.maxic, d.maxiv, d.maxido = self._memory_params(npar, npx, npy, npz, table)
        d.device_medium = 1 if self._device_setup else 0
        d.nstokes = self._nstokes
        d.nmicro = nmicro
        d.interpmethod = 1 if self._interpmethod == "OO" else 0
        self._keep = [ext, alb, iph, pw, zl, legenp]
        self._desc = d
        L = _lib.lib()
        h = C.c_void_p()
        err = _lib.errbuf()
        _lib.check(L.at3d_plan_create(C.byref(d), C.byref(h), err, len(err)), err)
        self._plan = h
        self._npts = int(L.at3d_plan_size(h, b"npts"))   # the current grid's (splitting adds points)
        self._nbpts = self._npts                          # the base grid's
        self._nlm = int(L.at3d_plan_size(h, b"nlm"))
        self._nang = int(L.at3d_plan_size(h, b"nang"))
        self._ml = int(L.at3d_plan_size(h, b"ml"))
        self._mm = int(L.at3d_plan_size(h, b"mm"))

    def _memory_params(self, npar, npx, npy, npz, table):
        """Array capacities of the adaptive grid as RTE._setup_memory sizes
        them (solver.py:2170-2322): MAXIG, MAXIC, MAXIV and MAXIDO from
        max_total_mb, adapt_grid_factor, cell_to_point_ratio and
        num_sh_term_factor; the splitting stops where these run out."""
        import psutil
        bc_open = npar["x_boundary_condition"] == "open"
        nx1 = npx if bc_open else npx + 1
        ny1 = npy if bc_open else npy + 1
        nbpts = nx1 * ny1 * npz
        nbcells = (npx + 1 if bc_open else npx) * (npy + 1 if bc_open else npy) * (npz - 1)
        nmu = max(2, 2 * ((int(npar["num_mu_bins"]) + 1) // 2))
        nphi = max(1, int(npar["num_phi_bins"]))
        ml, mm = nmu - 1, max(0, nphi // 2 - 1)
        nlm = (2 * mm + 1) * (ml + 1) - mm * (mm + 1)
        nphi0max = nphi
        max_total_mb = float(npar["max_total_mb"])
        if max_total_mb * 1024 ** 2 > 0.9 * psutil.virtual_memory().total:
            max_total_mb = 0.9 * psutil.virtual_memory().total / 1024 ** 2
        splitacc = float(npar["split_accuracy"])
        adapt = float(npar["adapt_grid_factor"])
        if splitacc < 0.0:
            adapt = 1.0
        big_arrays = 3 if bool(npar["acceleration_flag"]) else 2
        maxpg = npx * npy * npz
        numphase, maxleg = table.shape[1], table.shape[0]   # (the legendre_index size)
        nst = self._nstokes
        nstleg = 1 if nst == 1 else 6
        memword = nmu * (2 + 2 * nphi + 2 * nlm + 2 * 33 * 32) + 4.5 * maxpg + 2 * numphase * nstleg * maxleg
        ratio, shfac = float(npar["cell_to_point_ratio"]), float(npar["num_sh_term_factor"])
        wantmem = adapt * nbpts * (28 + 16.5 * ratio + nphi0max * nst + shfac * nst * nlm * big_arrays)
        reduce = min(1.0, ((max_total_mb * 1024 ** 2) / 4 - memword) / wantmem)
        adapt *= reduce
        if adapt < 1.0:
            raise _lib.SHDOMError(
                3, "max_total_mb memory limit ({}) exceeded with just base grid (final "
                "adapt_grid_factor={}). Either increase memory or reduce accuracy parameters "
                "(num_mu_bins={}, num_phi_bins={}, num_sh_term_factor={})".format(
                    max_total_mb, adapt, npar["num_mu_bins"], npar["num_phi_bins"], shfac))
        maxig = int(adapt * nbpts)
        maxic = max(int(ratio * maxig), nbcells)
        maxiv = max(int(shfac * nlm * maxig), nbpts * 4)
        return maxig, maxic, maxiv, maxig * nphi0max

    def _sync_grid(self):
        """The current grid's size after a solve (adaptive splitting adds points)."""
        if self._state is not None:
            self._npts = int(_lib.lib().at3d_state_size(self._state, b"npts"))

    def state_size(self, key: str) -> int:
        """Sizes of the solver state's current grid ("npts", "ncells", "nbpts",
        "nbcells", "ntop", "nbot", "refined")."""
        self._ensure_state()
        return int(_lib.lib().at3d_state_size(self._state, key.encode()))

    @property
    def final_adapt_grid_factor(self):
        """Points of the (refined) grid over the base grid's (solver.py:255)."""
        return None if self._state is None else self._npts / self._nbpts

    # -- plan inspection ---------------------------------------------------------
    def plan_size(self, key: str) -> int:
        return int(_lib.lib().at3d_plan_size(self._plan, key.encode()))

    def footprint_bytes(self) -> int:
        """Device bytes the solver state of this RTE allocates (known before
        it is allocated)."""
        return self.plan_size("footprint_bytes")

    _PLAN_SHAPES = {
        "gridpos": ("f4", lambda s: (s._nbpts, 3)), "dirflux": ("f4", lambda s: (s._nbpts,)),
        "total_ext": ("f4", lambda s: (s._nbpts,)), "radiance0": ("f4", lambda s: (s._nbpts, 4)),
        "ymat": ("f4", lambda s: (s._nang, s._nlm)), "wmat": ("f4", lambda s: (s._nang, s._nlm)),
        "ylmsun": ("f4", lambda s: (s._nlm,)), "mu": ("f4", lambda s: (s.plan_size("nmu"),)),
        "wtmu": ("f4", lambda s: (s.plan_size("nmu"),)), "nphi0": ("i4", lambda s: (s.plan_size("nmu"),)),
        "phi": ("f4", lambda s: (s.plan_size("nmu"), s.plan_size("nphi0max"))),
        "wtdo": ("f4", lambda s: (s.plan_size("nmu"), s.plan_size("nphi0max"))),
        "gridptr": ("i4", lambda s: (s.plan_size("ncells"), 8)),
        "neighptr": ("i4", lambda s: (s.plan_size("ncells"), 6)),
        "cellflags": ("i1", lambda s: (s.plan_size("ncells"),)),
        "sweepord": ("i4", lambda s: (8, s._nbpts)), "level": ("i4", lambda s: (8, s._nbpts)),
        "lvl_pts": ("i4", lambda s: (int((s.plan_array("level") > 0).sum()),)),
        "ring_level": ("i4", lambda s: (8, s._nbpts)), "ring_pos": ("i4", lambda s: (8, s._nbpts)),
        "ring_staged": ("i4", lambda s: (8, s._nbpts)), "ring_sizes": ("i4", lambda s: (10,)),
        "chi": ("f4", lambda s: (s._nbpts, s._ml + 1)), "alb": ("f4", lambda s: (s._nbpts,)),
        "extinct": ("f4", lambda s: (s.plan_size("npart"), s._nbpts)),
        "albedo": ("f4", lambda s: (s.plan_size("npart"), s._nbpts)),
        "bcptr_top": ("i4", lambda s: (s.plan_size("ntop"),)),
        "bcptr_bot": ("i4", lambda s: (s.plan_size("nbot"),)),
        "xgrid": ("f4", lambda s: (s.plan_size("nx1"),)), "ygrid": ("f4", lambda s: (s.plan_size("ny1"),)),
        "zgrid": ("f4", lambda s: (s.plan_size("nz"),)),
        "ypol": ("f4", lambda s: (6, s._nang, s._nlm)), "wpol": ("f4", lambda s: (6, s._nang, s._nlm)),
        "phasetab": ("f4", lambda s: (s.plan_size("numphase"), s.plan_size("nscatangle"))),
        "phasetab2": ("f4", lambda s: (s.plan_size("numphase"), s.plan_size("nscatangle"))),
        "legen": ("f4", lambda s: (s.plan_size("numphase"), s.plan_size("nleg") + 1)),
        "legen2": ("f4", lambda s: (s.plan_size("numphase"), s.plan_size("nleg") + 1)),
        "legen3": ("f4", lambda s: (s.plan_size("numphase"), s.plan_size("nleg") + 1)),
        "legen5": ("f4", lambda s: (s.plan_size("numphase"), s.plan_size("nleg") + 1)),
        "chip": ("f4", lambda s: (s._nbpts, 3, s._ml + 1)), "tms5": ("f4", lambda s: (s._nbpts, s._ml + 1)),
    }

    def plan_array(self, name: str) -> np.ndarray:
        dt, shp = self._PLAN_SHAPES[name]
        out = np.zeros(shp(self), dtype=dt)
        err = _lib.errbuf()
        _lib.check(_lib.lib().at3d_plan_get(self._plan, name.encode(), out.ctypes.data,
                                            out.nbytes, err, len(err)), err)
        return out

    # -- solve -------------------------------------------------------------------
    def _ensure_state(self):
        if self._state is None:
            h = C.c_void_p()
            err = _lib.errbuf()
            _lib.check(_lib.lib().at3d_state_create(self._plan, C.byref(h), err, len(err)), err)
            self._state = h

    def solve(self, maxiter, init_solution=True, setup_grid=True, verbose=True, solve=True):
        """Source-function iterations (solver.py:279-631)."""
        if not isinstance(verbose, bool):
            raise TypeError("`verbose` should be a boolean.")
        if not isinstance(init_solution, bool):
            raise TypeError("`init_solution` should be a boolean.")
        self._ensure_state()
        L = _lib.lib()
        err = _lib.errbuf()
        if init_solution or self._solcrit is None:
            if self._restore_data is not None:
                self._load_restore_data()
            else:
                _lib.check(L.at3d_init_solution(self._state, err, len(err)), err)
            self._solcrit, self._iters = 1.0, 0
            self._sync_grid()
        if not solve:
            return
        sc, it = C.c_float(), C.c_int32()
        _lib.check(L.at3d_solve(self._state, int(maxiter), int(verbose), C.byref(sc), C.byref(it),
                                err, len(err)), err)
        self._solcrit, self._iters = float(sc.value), int(it.value)
        self._sync_grid()
        if not self.check_solved(verbose=False):
            warnings.warn("solution did not converge to solution_accuracy={} in maxiter={} "
                          "(SOLCRIT={})".format(self.solution_accuracy, maxiter, self._solcrit))

    def init_solution_from(self, other):
        """Warm start from another RTE's current solution on the same grid and
        angular resolution: its spherical-harmonic radiance becomes this
        solver's initial radiance (instead of the two-stream start), and the
        next ``solve(..., init_solution=False)`` iterates from there."""
        if not isinstance(other, RTE) or other._state is None:
            raise ValueError("`other` must be an RTE holding a solution")
        self._ensure_state()
        err = _lib.errbuf()
        _lib.check(_lib.lib().at3d_init_solution_from(self._state, other._state, err, len(err)), err)
        self._solcrit, self._iters = 1.0, 0

    # -- saved solutions (solver.py:1519-1686) -----------------------------------
    def save_solution(self, save_radiances=True):
        """The grid and, with ``save_radiances``, the spherical-harmonic source
        and radiance of the current solution, under the variable names and
        layouts of the reference's saved-solution dataset (solver.py:1621-1686;
        Fortran order: gridpos (3, npts), gridptr (8, ncells) and neighptr
        (6, ncells) 1-based, source (nstokes, SHPTR[npts]), radiance
        (nstokes, RSHPTR[npts])).  A dict of numpy arrays (xarray is not used)."""
        g = self.grid_arrays()
        npts, ncells = g["gridpos"].shape[0], g["gridptr"].shape[0]
        # the 0-based pointers here map to the reference's 1-based ones by +1
        # (-1 = none -> 0; -(c+2), a pointer to split cell c -> -(c+1))
        out = {"nstokes": self._nstokes, "nx": self.plan_size("nx"), "ny": self.plan_size("ny"),
               "nz": self.plan_size("nz"), "ml": self._ml, "mm": self._mm, "nlm": self._nlm,
               "npts": npts, "ncells": ncells, "nbcells": self.plan_size("ncells"),
               "xgrid": self.plan_array("xgrid"), "ygrid": self.plan_array("ygrid"),
               "zgrid": self.plan_array("zgrid"),
               "gridpos": np.asfortranarray(g["gridpos"].T),
               "gridptr": np.asfortranarray(g["gridptr"].T + 1),
               "neighptr": np.asfortranarray(g["neighptr"].T + 1),
               "treeptr": np.asfortranarray(g["treeptr"].T + 1),
               "cellflags": g["cellflags"].astype(np.int16)}
        if save_radiances:
            if self._state is None or self._solcrit is None:
                raise RuntimeError("no solution to save: call solve() first")
            shptr = np.zeros(npts + 1, np.int32)
            rshptr = np.zeros(npts + 2, np.int32)
            fluxes = np.zeros((2, npts), np.float32)
            L, err = _lib.lib(), _lib.errbuf()
            _lib.check(L.at3d_state_save(self._state, shptr.ctypes.data, rshptr.ctypes.data, None, 0,
                                         None, 0, None, err, len(err)), err)
            # (nstokes, terms): the library writes the components one after another
            source = np.zeros((self._nstokes, int(shptr[npts])), np.float32)
            radiance = np.zeros((self._nstokes, int(rshptr[npts])), np.float32)
            _lib.check(L.at3d_state_save(self._state, shptr.ctypes.data, rshptr.ctypes.data,
                                         source.ctypes.data, source.size, radiance.ctypes.data,
                                         radiance.size, fluxes.ctypes.data, err, len(err)), err)
            out.update(fluxes=fluxes, shptr=shptr, rshptr=rshptr, source=source, radiance=radiance)
        return out

    def grid_arrays(self) -> dict:
        """The solver's current grid, refined by adaptive cell splitting after a
        solve: gridpos (npts, 3), gridptr (ncells, 8), neighptr (ncells, 6),
        treeptr (ncells, 2) and cellflags (ncells), 0-based (neighptr: c an end
        cell, -1 a boundary, -(c+2) a split cell c; treeptr: parent, first
        child, -1 none)."""
        if self._state is None:
            nc = self.plan_size("ncells")
            return {"gridpos": self.plan_array("gridpos"), "gridptr": self.plan_array("gridptr"),
                    "neighptr": self.plan_array("neighptr"),
                    "treeptr": np.full((nc, 2), -1, np.int32), "cellflags": self.plan_array("cellflags")}
        npts, nc = self.state_size("npts"), self.state_size("ncells")
        shapes = {"gridpos": ((npts, 3), np.float32), "gridptr": ((nc, 8), np.int32),
                  "neighptr": ((nc, 6), np.int32), "treeptr": ((nc, 2), np.int32),
                  "cellflags": ((nc,), np.int8)}
        out = {}
        L, err = _lib.lib(), _lib.errbuf()
        for name, (shape, dt) in shapes.items():
            a = np.zeros(shape, dt)
            _lib.check(L.at3d_state_get(self._state, name.encode(), a.ctypes.data, a.nbytes, err,
                                        len(err)), err)
            out[name] = a
        return out

    def load_solution(self, input_dataset, load_radiance=True):
        """Take the grid and spherical-harmonic solution of a saved solution
        (``save_solution`` here or the reference's dataset as a mapping of
        arrays, or an ``.npz`` path); the next ``solve(init_solution=True)``
        starts from it instead of the two-stream initialisation
        (solver.py:1519-1619, 2654-2664).  Raises ValueError for a grid or
        spherical-harmonic layout inconsistent with this solver."""
        ds = input_dataset
        if isinstance(ds, (str, bytes)) or hasattr(ds, "__fspath__"):
            with np.load(ds, allow_pickle=False) as f:
                ds = {k: f[k] for k in f.files}
        val = lambda k: np.asarray(ds[k]).item() if np.ndim(ds[k]) == 0 else np.asarray(ds[k])
        nx, ny, nz = self.plan_size("nx"), self.plan_size("ny"), self.plan_size("nz")
        if (val("nx"), val("ny"), val("nz")) != (nx, ny, nz):
            raise ValueError(
                "Incompatible grid sizes in loaded solution. self: (nx={}, ny={}, nz={}). "
                "!= loaded: (nx={}, ny={}, nz={}).".format(nx, ny, nz, val("nx"), val("ny"), val("nz")))
        for g in ("xgrid", "ygrid", "zgrid"):
            a, b = np.asarray(ds[g], np.float32), self.plan_array(g)
            if a.shape != b.shape or np.any(a != b):
                raise ValueError("Incompatible grids (xgrid, ygrid, zgrid) between loaded solution and self.")
        # the grid (possibly refined by adaptive splitting), 0-based; it becomes
        # the solver's grid when the solution is loaded (solver.py:1567-1595)
        npts, ncells = int(val("npts")), int(val("ncells"))
        if npts > self._desc.maxig or ncells > max(self._desc.maxic, self.plan_size("ncells")):
            raise _lib.SHDOMError(
                1, "Cannot load solution as loaded grid has more points ({}) or cells ({}) than "
                   "supported by the current RTE's memory parameters ({}, {})".format(
                       npts, ncells, self._desc.maxig, self._desc.maxic))
        grid = {"npts": npts, "ncells": ncells,
                "gridpos": np.ascontiguousarray(np.asarray(ds["gridpos"], np.float32)[:, :npts].T),
                "gridptr": np.ascontiguousarray(np.asarray(ds["gridptr"], np.int32)[:, :ncells].T - 1),
                "neighptr": np.ascontiguousarray(np.asarray(ds["neighptr"], np.int32)[:, :ncells].T - 1),
                "treeptr": np.ascontiguousarray(np.asarray(ds["treeptr"], np.int32)[:, :ncells].T - 1),
                "cellflags": np.ascontiguousarray(np.asarray(ds["cellflags"])[:ncells].astype(np.int8))}
        if not load_radiance:
            return
        if int(val("nstokes")) != self._nstokes:
            raise ValueError("Incompatible number of Stokes components in loaded solution. self: "
                             "nstokes={} != loaded: nstokes={}".format(self._nstokes, int(val("nstokes"))))
        if int(val("ml")) > self._ml or int(val("mm")) > self._mm:
            raise ValueError("Incompatible SphericalHarmonics in loaded solution. self: (ml={}, mm={}) "
                             "< loaded: (ml={}, mm={}).".format(self._ml, self._mm, int(val("ml")),
                                                               int(val("mm"))))
        if (int(val("ml")), int(val("mm"))) != (self._ml, self._mm):
            raise ValueError("a loaded solution must have this solver's (ml, mm) = ({}, {})"
                             .format(self._ml, self._mm))
        for name in ("radiance", "source", "fluxes", "shptr", "rshptr"):
            if name not in ds:
                raise KeyError("`input_dataset` does not appear to contain variable '{}' "
                               "which is required.".format(name))
        self._restore_data = {k: np.asarray(ds[k]) for k in ("radiance", "source", "fluxes", "shptr",
                                                             "rshptr")}
        self._restore_data["grid"] = grid

    def _load_restore_data(self):
        d, g = self._restore_data, self._restore_data["grid"]
        npts = g["npts"]
        err = _lib.errbuf()
        _lib.check(_lib.lib().at3d_state_set_grid(
            self._state, npts, g["ncells"], g["gridpos"].ctypes.data, g["gridptr"].ctypes.data,
            g["neighptr"].ctypes.data, g["treeptr"].ctypes.data, g["cellflags"].ctypes.data, err,
            len(err)), err)
        self._npts = npts
        shptr = _lib.as_i32(d["shptr"][:npts + 1])
        rshptr = np.zeros(npts + 2, np.int32)
        rshptr[:npts + 1] = d["rshptr"][:npts + 1]
        rshptr[npts + 1] = rshptr[npts]
        nst = self._nstokes
        src = np.asarray(d["source"]).reshape(nst, -1)
        rad = np.asarray(d["radiance"]).reshape(nst, -1)
        if src.shape[1] < shptr[npts] or rad.shape[1] < rshptr[npts]:
            raise ValueError("loaded source / radiance hold fewer terms than their pointers give")
        # the Stokes components one after another
        source = _lib.as_f32(src[:, :int(shptr[npts])].ravel())
        radiance = _lib.as_f32(rad[:, :int(rshptr[npts])].ravel())
        fluxes = _lib.as_f32(np.asarray(d["fluxes"])[:, :npts])
        err = _lib.errbuf()
        _lib.check(_lib.lib().at3d_state_load(self._state, shptr.ctypes.data, source.ctypes.data,
                                              rshptr.ctypes.data, radiance.ctypes.data,
                                              fluxes.ctypes.data, err, len(err)), err)

    # -- diagnostics on the base grid (solver.py:1037-1262) -----------------------
    def _base_grid(self, arr):
        return arr[..., :self._nbpts].reshape(arr.shape[:-1] + (self.plan_size("nx1"),
                                                                self.plan_size("ny1"),
                                                                self.plan_size("nz")))

    def _coords(self):
        return {"x": self.plan_array("xgrid"), "y": self.plan_array("ygrid"),
                "z": self.plan_array("zgrid")}

    def _sh_terms(self):
        nshout = 5 if bool(self.numerical_params["high_order_radiance"]) else 4
        sh = np.zeros((nshout, self._npts), np.float32)
        nfd = np.zeros(self._npts, np.float32)
        err = _lib.errbuf()
        _lib.check(_lib.lib().at3d_state_sh_terms(self._state, nshout, sh.ctypes.data, nfd.ctypes.data,
                                                  err, len(err)), err)
        return sh, nfd

    @property
    def adaptive_shterms(self):
        """Mean radiance and net fluxes (Fx, Fy, Fz) at every grid point
        (COMPUTE_SH), [nshout, npts]."""
        self._ensure_state()
        return self._sh_terms()[0]

    @property
    def spherical_harmonics(self):
        """Mean intensity and net fluxes Fx, Fy, Fz on the base grid, plus
        'rms_higher_rad' with high_order_radiance (solver.py:1037-1121)."""
        self.check_solved()
        sh, _ = self._sh_terms()
        out = {"mean_intensity": self._base_grid(sh[0]), "Fx": self._base_grid(sh[1]),
               "Fy": self._base_grid(sh[2]), "Fz": self._base_grid(sh[3])}
        if sh.shape[0] == 5:
            out["rms_higher_rad"] = self._base_grid(sh[4]) / np.sqrt(np.pi * 4.0 * self._nlm)
        out.update(self._coords())
        return out

    @property
    def adaptive_fluxes(self):
        return self.state_array("fluxes")

    @property
    def fluxes(self):
        """Down- and upwelling hemispherical fluxes and the direct solar flux on
        the base grid (solver.py:1148-1198)."""
        self.check_solved()
        f = self.state_array("fluxes")
        out = {"flux_down": self._base_grid(f[0]), "flux_up": self._base_grid(f[1]),
               "flux_direct": self._base_grid(self.state_array("dirflux"))}
        out.update(self._coords())
        return out

    @property
    def adaptive_net_flux_div(self):
        self._ensure_state()
        return self._sh_terms()[1]

    @property
    def net_flux_divergence(self):
        """Net flux divergence on the base grid (solver.py:1207-1262)."""
        self.check_solved()
        out = {"net_flux_div": self._base_grid(self._sh_terms()[1])}
        out.update(self._coords())
        return out

    def integrate_to_sensor(self, sensor, single_scatter=False, nosurface=False):
        """Radiance of every ray of ``sensor`` (solver.py:633-791).

        ``sensor`` maps 'ray_x', 'ray_y', 'ray_z' (km), 'ray_mu' (cosine zenith of
        the direction towards the sensor) and 'ray_phi' (radians) to 1-D arrays;
        'I' is added to it and the sensor is returned."""
        for k in ("ray_x", "ray_y", "ray_z", "ray_mu", "ray_phi"):
            if k not in sensor:
                raise KeyError(f"sensor is missing '{k}'")
        requested = [str(v) for v in np.atleast_1d(sensor.get("stokes", ["I"]))]
        if len(requested) > self._nstokes:
            raise ValueError("'{}' Stokes components are requested by the sensor but the solver "
                             "has nstokes={}".format(len(requested), self._nstokes))
        self.check_solved()
        self._ensure_state()
        x = _lib.as_f32(sensor["ray_x"])
        y = _lib.as_f32(sensor["ray_y"])
        z = _lib.as_f32(sensor["ray_z"])
        mu = np.ascontiguousarray(sensor["ray_mu"], dtype=np.float64)
        phi = np.ascontiguousarray(sensor["ray_phi"], dtype=np.float64)
        n = x.size
        if not all(a.size == n for a in (y, z, mu, phi)):
            raise ValueError("ray arrays must have the same length")
        out = np.zeros((n, self._nstokes), np.float32)
        self.set_option("render_correct_interpolate", 1 if self._correctinterpolate else 0)
        err = _lib.errbuf()
        _lib.check(_lib.lib().at3d_render(self._state, n, x.ctypes.data, y.ctypes.data,
                                          z.ctypes.data, mu.ctypes.data, phi.ctypes.data,
                                          int(single_scatter), int(nosurface), out.ctypes.data,
                                          err, len(err)), err)
        sensor["I"] = np.ascontiguousarray(out[:, 0])
        if self._nstokes > 1:
            # Stokes parameters for linear polarization (solver.py:768-781)
            sensor["Q"] = np.ascontiguousarray(out[:, 1])
            sensor["U"] = np.ascontiguousarray(out[:, 2])
        return sensor

    def check_solved(self, verbose=True):
        """True once SOLCRIT <= solution_accuracy (solver.py:1012-1035)."""
        ok = self._solcrit is not None and self._solcrit <= self.solution_accuracy
        if not ok and verbose:
            warnings.warn("RTE is not solved to the requested accuracy")
        return ok

    @property
    def solution_accuracy(self):
        return float(self.numerical_params["solution_accuracy"])

    @property
    def num_iterations(self):
        return self._iters

    @property
    def solcrit(self):
        return self._solcrit

    def state_array(self, name: str) -> np.ndarray:
        """Copy a device array: source, radiance, delsource [npts, nlm] (polarized:
        [3, npts, nlm], the I, Q and U terms); fluxes [2, npts];
        dirflux [npts]; ns, nr [npts]; do_radiance, do_source [nang, npts]."""
        if self._state is None:
            raise RuntimeError("solve() has not been called")
        n, ml1 = self._npts, self._ml + 1
        sh = (n, self._nlm) if self._nstokes == 1 else (self._nstokes, n, self._nlm)
        if name in ("point_ss_iph", "point_ss_coef"):
            off = self.state_array("point_ss_off")
            shp, dt = (int(off[-1]),), (np.int32 if name == "point_ss_iph" else np.float32)
            out = np.zeros(shp, dt)
            err = _lib.errbuf()
            _lib.check(_lib.lib().at3d_state_get(self._state, name.encode(), out.ctypes.data,
                                                 out.nbytes, err, len(err)), err)
            return out
        shapes = {"point_ext": ((n,), np.float32), "point_alb": ((n,), np.float32),
                  "point_chi": ((n, ml1), np.float32), "point_tc": ((n, ml1), np.float32),
                  "point_tms": ((n, ml1), np.float32), "point_radiance0": ((n, 4), np.float32),
                  "point_ss_off": ((n + 1,), np.int32),
                  "source": (sh, np.float32), "radiance": (sh, np.float32), "delsource": (sh, np.float32),
                  "fluxes": ((2, self._npts), np.float32), "dirflux": ((self._npts,), np.float32),
                  "ns": ((self._npts,), np.int32), "nr": ((self._npts,), np.int32),
                  "do_radiance": ((self._nang, self._npts), np.float32),
                  "do_source": ((self._nang, self._npts), np.float32)}
        shp, dt = shapes[name]
        out = np.zeros(shp, dt)
        err = _lib.errbuf()
        _lib.check(_lib.lib().at3d_state_get(self._state, name.encode(), out.ctypes.data,
                                             out.nbytes, err, len(err)), err)
        return out

    def set_option(self, key: str, value: int):
        """Engine option of the device state (see at3d_state_set_option)."""
        self._ensure_state()
        err = _lib.errbuf()
        _lib.check(_lib.lib().at3d_state_set_option(self._state, key.encode(), int(value), err,
                                                    len(err)), err)

    def timing(self, key: str) -> float:
        if self._state is None:
            return -1.0
        return float(_lib.lib().at3d_state_timing(self._state, key.encode()))

    @property
    def _source(self):
        return self.state_array("source")

    @property
    def _radiance(self):
        return self.state_array("radiance")

    @property
    def _fluxes(self):
        return self.state_array("fluxes")

    @property
    def _dirflux(self):
        return self.plan_array("dirflux")

    def release(self):
        """Free device memory (the reference's _release_big_arrays, solver.py:2529)."""
        if self._state is not None:
            _lib.lib().at3d_state_destroy(self._state)
            self._state = None

    def __del__(self):
        try:
            self.release()
            if self._plan is not None:
                _lib.lib().at3d_plan_destroy(self._plan)
                self._plan = None
        except Exception:
            pass
