"""Seeded synthetic media for the benchmark and parity workloads.

Every generator returns plain numpy arrays (property grid = RTE base grid) so
that the same inputs can drive this engine and any checker.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Workload:
    name: str
    delx: float
    dely: float
    z: np.ndarray
    # list of (ext, ssalb, legcoef[(6,) nleg+1, ntab]) or (ext, ssalb, legcoef,
    # table_index, phase_weights) with 1-based table_index / phase_weights of
    # shape ([num_micro,] nx, ny, nz)
    species: list = field(default_factory=list)
    bc: str = "open"
    nmu: int = 16
    nphi: int = 32
    solacc: float = 1e-4
    maxiter: int = 100
    solarmu: float = -0.5
    solaraz_deg: float = 0.0
    gndalbedo: float = 0.05
    wavelength: float = 0.86
    deltam: bool = True
    nstokes: int = 1    # 1, or 3 for a polarized solve (legcoef then (6, nleg+1, ntab))
    interpmethod: str = "ON"   # property transfer: 'ON' mixes tables, 'OO' takes one

    @property
    def shape(self):
        return self.species[0][0].shape


def hg_legendre(g: float, nleg: int) -> np.ndarray:
    """Henyey-Greenstein Legendre coefficients (2l+1) g^l, l = 0..nleg, as a one-table column."""
    l = np.arange(nleg + 1, dtype=np.float64)
    return ((2 * l + 1) * g ** l)[:, None]


def lognormal_blob(n=16, dx=0.05, dz=0.04, radius=0.3, mean=10.0, sigma=0.5, seed=0,
                   ssalb=0.999, g=0.85, nleg=7, bc="periodic", nmu=8, nphi=16, solacc=1e-5):
    """Config-1 medium: lognormal extinction (mean `mean`) inside a sphere of
    radius `radius` km about the domain centre, zero outside; HG phase
    function with `nleg`+1 Legendre terms."""
    rng = np.random.default_rng(seed)
    x = np.arange(n) * dx
    z = (np.arange(n) * dz).astype(np.float32)
    X, Y, Z = np.meshgrid(x, x, z, indexing="ij")
    c = np.array([x[-1] / 2 + dx / 2 if bc == "periodic" else x[-1] / 2, x[-1] / 2, z[-1] / 2])
    if bc == "periodic":
        c[1] = c[0]
    r = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)
    inside = r < radius
    ext = np.zeros((n, n, n), np.float32)
    ext[inside] = mean * np.exp(sigma * rng.standard_normal(inside.sum()) - 0.5 * sigma ** 2)
    alb = np.full(ext.shape, ssalb, np.float32)
    return Workload("lognormal_blob", dx, dx, z, [(ext, alb, hg_legendre(g, nleg))], bc=bc,
                    nmu=nmu, nphi=nphi, solacc=solacc)


def gaussian_field(shape, slope=-5.0 / 3.0, seed=0):
    """Isotropic Gaussian random field with power spectrum |k|^slope (normalised to
    zero mean, unit variance): white noise filtered by |k|^(slope/2) in Fourier space."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape)
    ks = np.meshgrid(*[np.fft.fftfreq(s) for s in shape], indexing="ij")
    k = np.sqrt(sum(kk ** 2 for kk in ks))
    k[(0,) * len(shape)] = np.inf
    f = np.real(np.fft.ifftn(np.fft.fftn(noise) * k ** (slope / 2.0)))
    return (f - f.mean()) / f.std()


def cloud_field(nx=64, ny=64, nz=40, dx=0.02, dz=0.025, z0=0.0, seed=0, cover=0.4,
                ext_scale=30.0, cloud_base=0.5, cloud_top=1.2, rayleigh=0.01,
                ssalb=0.9995, g=0.86, nleg_table=199, nmu=16, nphi=32, solacc=1e-4):
    """Config-2 medium: a spectral-slope -5/3 Gaussian field thresholded at
    (1-cover) quantile, extinction ext_scale * field**1.5 (field shifted to start
    at 0 above the threshold) between cloud_base and cloud_top km, plus a
    Rayleigh species of 0.01/km with albedo 1.  The 40 levels at 0.025 km span
    0.975 km, so the vertical grid is placed at z0 + k*dz with z0 = 0.35 km, which
    puts the 0.5-1.2 km cloud layer inside the domain."""
    z = (z0 + np.arange(nz) * dz).astype(np.float32)
    f = gaussian_field((nx, ny, nz), seed=seed)
    thr = np.quantile(f, 1.0 - cover)
    cl = np.clip(f - thr, 0.0, None)
    cl = cl / max(cl.max(), 1e-12)
    layer = (z >= cloud_base) & (z <= cloud_top)
    ext = (ext_scale * cl ** 1.5 * layer[None, None, :]).astype(np.float32)
    alb = np.full(ext.shape, ssalb, np.float32)
    l = np.arange(nleg_table + 1, dtype=np.float64)
    cloud_leg = (np.exp(-l / 30.0) * (2 * l + 1) * g ** l)
    cloud_leg[0] = 1.0
    rext = np.full(ext.shape, rayleigh, np.float32)
    ralb = np.ones(ext.shape, np.float32)
    rleg = np.zeros(3)
    rleg[0], rleg[2] = 1.0, 0.5   # Rayleigh phase function: (2l+1) chi_l with chi_2 = 0.1
    return Workload("cloud_field", dx, dx, z,
                    [(ext, alb, cloud_leg[:, None]), (rext, ralb, rleg[:, None])],
                    bc="open", nmu=nmu, nphi=nphi, solacc=solacc)


def rayleigh_phase_matrix(depol=0.0, nleg=2):
    """The six Legendre expansions (a1, a2, a3, a4, b1, b2, each including
    the 2l+1 factor) of the Rayleigh phase matrix with depolarization factor
    `depol`: a1 = (1, 0, D/2), a2_2 = 3D, a4_1 = 3 D D'/2, b1_2 = sqrt(6) D/2
    with D = (1 - depol)/(1 + depol/2) and D' = (1 - 2 depol)/(1 - depol)
    (Hovenier, van der Mee and Domke, 2004).  Returns (6, nleg+1, 1)."""
    dd = (1.0 - depol) / (1.0 + 0.5 * depol)
    dp = (1.0 - 2.0 * depol) / (1.0 - depol)
    t = np.zeros((6, max(nleg, 2) + 1, 1))
    t[0, 0, 0] = 1.0
    t[0, 2, 0] = 0.5 * dd
    t[1, 2, 0] = 3.0 * dd
    t[3, 1, 0] = 1.5 * dd * dp
    t[4, 2, 0] = np.sqrt(6.0) * 0.5 * dd
    return t


def polarized_cloud_table(g=0.85, nleg=199, seed=0):
    """A smooth decaying six-element phase matrix expansion (Henyey-Greenstein
    a1 damped by exp(-l/30); a2 = a3 and a4 proportional to it from l = 2 and
    l = 1; b1, b2 negative and decaying faster), with seeded proportions.
    Returns (6, nleg+1, 1)."""
    rng = np.random.default_rng(seed)
    l = np.arange(nleg + 1, dtype=np.float64)
    a1 = np.exp(-l / 30.0) * (2 * l + 1) * g ** l
    a1[0] = 1.0
    c2, c3, c4 = rng.uniform(0.85, 0.95), rng.uniform(0.8, 0.9), rng.uniform(0.75, 0.85)
    cb1, cb2 = rng.uniform(0.05, 0.15), rng.uniform(0.0, 0.05)
    t = np.zeros((6, nleg + 1, 1))
    t[0, :, 0] = a1
    t[1, 2:, 0] = c2 * a1[2:]
    t[2, 2:, 0] = c3 * a1[2:]
    t[3, 1:, 0] = c4 * a1[1:]
    t[4, 2:, 0] = -cb1 * a1[2:] * np.exp(-(l[2:] - 2) / 10.0)
    t[5, 2:, 0] = -cb2 * a1[2:] * np.exp(-(l[2:] - 2) / 10.0)
    return t


def polarized_cloud_field(nx=128, ny=128, nz=64, dx=0.02, dz=0.02, z0=0.0, seed=0, wavelength=0.66,
                          ext_scale=30.0, cloud_base=0.3, cloud_top=1.0, ssalb=0.9995, g=0.85,
                          nleg_table=199, depol=0.03, nmu=16, nphi=32, solacc=1e-4):
    """Polarized medium (NSTOKES = 3): the cloud of cloud_field on this grid
    with a six-element table (polarized_cloud_table) and its extinction
    scaled by (0.86 / wavelength), plus Rayleigh scattering with
    depolarization `depol` whose extinction 0.0116 / km at 0.86 um scales as
    wavelength^-4.  The grid spans z0 .. z0 + (nz - 1) dz km; the cloud
    layer is clipped to it."""
    z = (z0 + np.arange(nz) * dz).astype(np.float32)
    f = gaussian_field((nx, ny, nz), seed=seed)
    thr = np.quantile(f, 0.6)
    cl = np.clip(f - thr, 0.0, None)
    cl = cl / max(cl.max(), 1e-12)
    layer = (z >= cloud_base) & (z <= cloud_top)
    scale = 0.86 / wavelength
    ext = (ext_scale * scale * cl ** 1.5 * layer[None, None, :]).astype(np.float32)
    alb = np.full(ext.shape, ssalb, np.float32)
    rext = np.full(ext.shape, 0.0116 * (0.86 / wavelength) ** 4, np.float32)
    ralb = np.ones(ext.shape, np.float32)
    return Workload("polarized_cloud", dx, dx, z,
                    [(ext, alb, polarized_cloud_table(g, nleg_table, seed)),
                     (rext, ralb, rayleigh_phase_matrix(depol))],
                    bc="open", nmu=nmu, nphi=nphi, solacc=solacc, wavelength=wavelength, nstokes=3)


def workload(name: str) -> Workload:
    if name == "blob16":
        w = lognormal_blob(n=16, nleg=64)
    elif name == "blob16_l7":
        w = lognormal_blob(n=16, nleg=7)
    elif name == "cloud64":
        w = cloud_field(z0=0.35)
    elif name == "polarized128":
        w = polarized_cloud_field()
    else:
        raise KeyError(name)
    return w


def make_rte(w: Workload, device_setup=True, **overrides):
    """An RTE for a workload (solar source, fixed Lambertian surface)."""
    from . import configuration, source, surface
    from .medium import Scatterer
    from .solver import RTE
    # the workloads are defined on fixed base grids (no adaptive splitting)
    cfg = configuration.get_config(num_mu_bins=w.nmu, num_phi_bins=w.nphi,
                                   x_boundary_condition=w.bc, y_boundary_condition=w.bc,
                                   solution_accuracy=w.solacc, deltam=w.deltam,
                                   split_accuracy=0.0)
    cfg.update(overrides)
    med = {f"species_{i}": Scatterer(sp[0], sp[1], sp[2], w.delx, w.dely, w.z, *sp[3:5])
           for i, sp in enumerate(w.species)}
    return RTE(cfg, med, source.solar(w.wavelength, w.solarmu, w.solaraz_deg),
               surface.lambertian(w.gndalbedo), num_stokes=w.nstokes, device_setup=device_setup,
               interpmethod=w.interpmethod)


# Serving stream: grids, (NMU, NPHI) angle sets (L = NMU - 1) and wavelengths
SERVE_GRIDS = ((32, 32, 32), (48, 48, 32), (64, 64, 40), (96, 96, 48))
SERVE_ANGLES = ((8, 16), (12, 24), (16, 32))


def serve_medium(grid, nmu, nphi, wavelength, seed, noise=None, noise_seed=0, solacc=1e-5):
    """One medium of the serving stream: the config-1 lognormal blob (mean
    10/km at 0.86 um, sigma 0.5, radius 0.3 km, HG g = 0.85 as 64 Legendre
    terms (the solve keeps L + 1 = NMU of them; a shorter table is negative at
    back-scatter, which rendering rejects), albedo 0.999) on an open grid of
    `grid` points over a 0.8 x 0.8 x
    0.64 km domain; extinction scales as (0.86 / wavelength).  `noise` (a
    relative sigma) multiplies the extinction by seeded lognormal noise, as an
    inversion step perturbs the medium."""
    nx, ny, nz = grid
    dx, dz = 0.8 / nx, 0.64 / nz
    rng = np.random.default_rng(seed)
    x = np.arange(nx) * dx
    y = np.arange(ny) * dx
    z = (np.arange(nz) * dz).astype(np.float32)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    r = np.sqrt((X - x[-1] / 2) ** 2 + (Y - y[-1] / 2) ** 2 + (Z - z[-1] / 2) ** 2)
    inside = r < 0.3
    ext = np.zeros((nx, ny, nz), np.float32)
    ext[inside] = 10.0 * (0.86 / wavelength) * np.exp(0.5 * rng.standard_normal(inside.sum()) - 0.125)
    if noise:
        nrng = np.random.default_rng(noise_seed)
        ext *= np.exp(noise * nrng.standard_normal(ext.shape) - 0.5 * noise ** 2).astype(np.float32)
    alb = np.full(ext.shape, 0.999, np.float32)
    return Workload("serve", dx, dx, z, [(ext, alb, hg_legendre(0.85, 64))], bc="open",
                    nmu=nmu, nphi=nphi, solacc=solacc, wavelength=float(wavelength))


def serving_stream(n_requests=512, rate_hz=20.0, warm_fraction=0.3, seed=0, grids=SERVE_GRIDS,
                   angles=SERVE_ANGLES, max_wavelengths=4, solacc=1e-5):
    """The serving stream's requests: Poisson arrivals at `rate_hz`; each
    request draws a grid, an angle set and 1..max_wavelengths wavelengths in
    0.55-2.2 um, and is one solve per wavelength (a ``Request`` each, sharing
    the request's ``group``).  A `warm_fraction` of requests restart from an
    earlier request (same grid, angles and wavelengths; extinction perturbed
    by 5 % lognormal noise) and seed each solve from that request's solution.
    Returns the list of ``serving.Request``."""
    from .serving import Request
    rng = np.random.default_rng(seed)
    t = 0.0
    specs, reqs = [], []
    for i in range(n_requests):
        t += rng.exponential(1.0 / rate_hz)
        src = None
        if specs and rng.random() < warm_fraction:
            src = int(rng.integers(len(specs)))
        if src is None:
            grid = grids[int(rng.integers(len(grids)))]
            nmu, nphi = angles[int(rng.integers(len(angles)))]
            nwl = int(rng.integers(1, max_wavelengths + 1))
            wls = np.sort(rng.uniform(0.55, 2.2, nwl)).round(4).tolist()
            mseed = int(rng.integers(1 << 31))
            spec = dict(grid=grid, nmu=nmu, nphi=nphi, wls=wls, seed=mseed, noise=None, src=None)
        else:
            s = specs[src]
            spec = dict(s, noise=0.05, noise_seed=int(rng.integers(1 << 31)), src=src)
        specs.append(spec)
        for j, wl in enumerate(spec["wls"]):
            def make(spec=spec, wl=wl):
                w = serve_medium(spec["grid"], spec["nmu"], spec["nphi"], wl, spec["seed"],
                                 spec["noise"], spec.get("noise_seed", 0), solacc=solacc)
                return make_rte(w)
            make.geometry = (tuple(spec["grid"]), spec["nmu"], spec["nphi"])
            reqs.append(Request(key=(i, j), make_rte=make, maxiter=100,
                                warm_from=None if spec["src"] is None else (spec["src"], j),
                                arrival_s=t, group=i))
    return reqs


def warm_geometries(requests):
    """Builds one plan for each distinct (grid, NMU, NPHI) of `requests`
    (made by ``serving_stream``), so the host's cache of sweep levels (shared
    by plans of one geometry) holds them all, as in a server that has seen
    its usual geometries.  Returns the number of geometries."""
    seen = {}
    for r in requests:
        g = getattr(r.make_rte, "geometry", None)
        if g is not None and g not in seen:
            seen[g] = r.make_rte
    for make in seen.values():
        make()
    return len(seen)


def domain_box(w: Workload):
    """(xmin, ymin, zmin, xmax, ymax, zmax) of a workload's base grid; a
    periodic grid spans one more cell in x and y than its points."""
    nx, ny, _ = w.shape
    extra = 1 if w.bc == "periodic" else 0
    return (0.0, 0.0, float(w.z[0]), (nx - 1 + extra) * w.delx, (ny - 1 + extra) * w.dely,
            float(w.z[-1]))


def view_rays(w: Workload, views, npix_x, npix_y):
    """Parallel-ray views of the domain from its top: for each (zenith,
    azimuth) in degrees (the direction of the photons), npix_x x npix_y
    pixels spread evenly over the footprint of the domain box on the top
    plane (the area orthographic_projection samples), x fastest.  Returns
    the concatenated 'ray_*' arrays."""
    xmin, ymin, zmin, xmax, ymax, zmax = domain_box(w)
    out = {k: [] for k in ("ray_x", "ray_y", "ray_z", "ray_mu", "ray_phi")}
    for zen, az in views:
        mu = float(np.cos(np.deg2rad(zen)))
        phi = float(np.deg2rad(az))
        t = np.sqrt(1.0 - mu * mu) / mu
        sx = t * np.cos(phi) * np.array([0.0, zmax - zmin])
        sy = t * np.sin(phi) * np.array([0.0, zmax - zmin])
        xs = np.linspace(xmin + sx.min(), xmax + sx.max(), npix_x)
        ys = np.linspace(ymin + sy.min(), ymax + sy.max(), npix_y)
        n = npix_x * npix_y
        out["ray_x"].append(np.tile(xs, npix_y))
        out["ray_y"].append(np.repeat(ys, npix_x))
        out["ray_z"].append(np.full(n, zmax))
        out["ray_mu"].append(np.full(n, mu))
        out["ray_phi"].append(np.full(n, phi))
    return {k: np.concatenate(v) for k, v in out.items()}


# the headline workload's sensors: 9 views of 64 x 64 pixels along the
# azimuth plane 0 / 180 degrees; the small blob's: 3 views of 32 x 32
CLOUD64_VIEWS = tuple((abs(z), 0.0 if z >= 0 else 180.0) for z in (-60, -45, -30, -15, 0, 15, 30, 45, 60))
BLOB16_VIEWS = ((0.0, 0.0), (45.0, 0.0), (45.0, 180.0))


def workload_views(name: str, w: Workload):
    if name == "cloud64":
        return view_rays(w, CLOUD64_VIEWS, 64, 64)
    if name == "polarized128":
        return view_rays(w, CLOUD64_VIEWS, 128, 128)
    if name.startswith("blob16"):
        return view_rays(w, BLOB16_VIEWS, 32, 32)
    raise KeyError(name)
