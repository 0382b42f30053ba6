"""Small seeded scenes shared by the parity tests."""
import numpy as np

from at3d_amd import synthetic as S
from at3d_amd.sensor import orthographic_projection


def blob8(bc="periodic", nleg=64, seed=0, nmu=8, nphi=16):
    return S.lognormal_blob(n=8, radius=0.15, seed=seed, nleg=nleg, bc=bc, nmu=nmu, nphi=nphi)


def blob8_l15(bc="open"):
    """The 8^3 blob at the headline angular resolution: NMU 16 / NPHI 32,
    L = M = 15 (NLM 256, NANG 354), 64-term HG table (delta-M f = chi_16)."""
    return blob8(bc, nmu=16, nphi=32)


def cloud_small(seed=1, n=12, nz=10):
    """Two species (cloud + Rayleigh) on an open grid, NMU 8 / NPHI 16."""
    w = S.cloud_field(nx=n, ny=n, nz=nz, dx=0.05, dz=0.1, z0=0.4, seed=seed, ext_scale=20.0,
                      nmu=8, nphi=16, solacc=1e-5)
    return w


def cloud_l15(seed=2, n=16, nz=12):
    """Two species at the headline angular resolution (NMU 16 / NPHI 32,
    L = 15): the cloud with the 200-term phase table (delta-M f = chi_16,
    single-scatter correction from the whole table) plus Rayleigh, on a
    16 x 16 x 12 open grid."""
    return S.cloud_field(nx=n, ny=n, nz=nz, dx=0.05, dz=0.08, z0=0.3, seed=seed, ext_scale=20.0,
                         nmu=16, nphi=32, solacc=1e-5)


def pol_blob(nmu=8, nphi=16, bc="open", n=8, seed=0):
    """Polarized (NSTOKES = 3) two-species scene: the 8^3 lognormal blob with
    a six-element cloud phase matrix (64 orders) plus Rayleigh scattering
    (depolarization 0.03, 0.05 / km) everywhere."""
    w = S.lognormal_blob(n=n, radius=0.15, seed=seed, nleg=64, bc=bc, nmu=nmu, nphi=nphi)
    e, a, _ = w.species[0]
    rext = np.full(e.shape, 0.05, np.float32)
    ralb = np.ones(e.shape, np.float32)
    species = [(e, a, S.polarized_cloud_table(0.85, 64, seed=seed + 1)), (rext, ralb, S.rayleigh_phase_matrix(0.03))]
    return S.Workload("pol_blob", w.delx, w.dely, w.z, species, bc=bc, nmu=nmu, nphi=nphi, solacc=1e-5,
                      nstokes=3)


def micro_blob(bc="open", nmicro=2, seed=0, nmu=8, nphi=16):
    """Several phase tables per property point (num_micro): the 8^3 blob with
    three HG tables (g = 0.8, 0.85, 0.9, 64 orders) mixed at every point by
    `nmicro` seeded (table, weight) pairs, plus Rayleigh scattering with one
    table (padded to `nmicro` with weight 0)."""
    w = blob8(bc, seed=seed, nmu=nmu, nphi=nphi)
    e, a, _ = w.species[0]
    rng = np.random.default_rng(seed + 7)
    tables = np.concatenate([S.hg_legendre(g, 64) for g in (0.8, 0.85, 0.9)], axis=1)
    ti = rng.integers(1, 4, size=(nmicro,) + e.shape).astype(np.int32)
    wt = rng.uniform(0.2, 1.0, size=(nmicro,) + e.shape)
    wt = (wt / wt.sum(axis=0)).astype(np.float32)
    rext = np.full(e.shape, 0.05, np.float32)
    ralb = np.ones(e.shape, np.float32)
    rleg = np.array([[1.0], [0.0], [0.5]])
    return S.Workload("micro_blob", w.delx, w.dely, w.z, [(e, a, tables, ti, wt), (rext, ralb, rleg)], bc=bc,
                      nmu=nmu, nphi=nphi, solacc=1e-5)


def oo_blob(bc="open", nstokes=1, seed=0, nmu=8, nphi=16):
    """INTERPMETHOD 'OO': the 8^3 blob with one of three tables per property
    point (seeded), HG g = 0.8 / 0.85 / 0.9 (64 orders) or, polarized, three
    six-element cloud tables; every grid point takes the table of the
    property point with the most scattering among its eight."""
    w = blob8(bc, seed=seed, nmu=nmu, nphi=nphi)
    e, a, _ = w.species[0]
    rng = np.random.default_rng(seed + 11)
    if nstokes == 1:
        tables = np.concatenate([S.hg_legendre(g, 64) for g in (0.8, 0.85, 0.9)], axis=1)
    else:
        tables = np.concatenate([S.polarized_cloud_table(g, 64, seed=seed + k)
                                 for k, g in enumerate((0.8, 0.85, 0.9))], axis=2)
    ti = rng.integers(1, 4, size=e.shape).astype(np.int32)
    return S.Workload("oo_blob", w.delx, w.dely, w.z, [(e, a, tables, ti)], bc=bc, nmu=nmu, nphi=nphi,
                      solacc=1e-5, nstokes=nstokes, interpmethod="OO")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def absorber_columns(nx=8, ny=4, nz=6, taumax=2.0, gndalbedo=0.04, bc="open"):
    """Pure absorber (single-scattering albedo 0) whose column optical depth
    grows linearly with x and is uniform in y and z; overhead sun, flux 1.

    Closed form at the column of x index i for a nadir view from the top:
    I = exp(-tau_i) * gndalbedo / pi * exp(-tau_i), tau_i = taumax * i / (nx - 1)."""
    z = np.linspace(0.0, 1.0, nz).astype(np.float32)
    tau = taumax * np.arange(nx) / (nx - 1)
    ext = np.broadcast_to((tau / (z[-1] - z[0]))[:, None, None], (nx, ny, nz)).astype(np.float32)
    alb = np.zeros((nx, ny, nz), np.float32)
    w = S.Workload("absorber_columns", 0.1, 0.1, z, [(np.ascontiguousarray(ext), alb,
                                                        S.hg_legendre(0.0, 0))], bc=bc,
                   nmu=8, nphi=16, solacc=1e-5, solarmu=-1.0, gndalbedo=gndalbedo)
    return w, tau


def bbox(w):
    nx, ny, _ = w.shape
    return (0.0, 0.0, float(w.z[0]), (nx - 1) * w.delx, (ny - 1) * w.dely, float(w.z[-1]))


def ortho_rays(w, views):
    xs, ys, zs, mus, phis = [], [], [], [], []
    for zen, az in views:
        s = orthographic_projection(0.86, bbox(w), w.delx * 0.7, w.dely * 0.7, az, zen)
        for k, acc in (("ray_x", xs), ("ray_y", ys), ("ray_z", zs), ("ray_mu", mus),
                       ("ray_phi", phis)):
            acc.append(s[k])
    return {k: np.concatenate(v) for k, v in zip(("ray_x", "ray_y", "ray_z", "ray_mu", "ray_phi"),
                                                 (xs, ys, zs, mus, phis))}


def perspective_rays(w, nx=8, ny=6, fov=40.0, height=0.0, sub=None):
    """A pinhole camera on the domain's top surface (plus `height`), above its
    centre and offset in x, looking at the domain's centre."""
    from at3d_amd.sensor import perspective_projection
    x0, y0, z0, x1, y1, z1 = bbox(w)
    pos = (x0 + 0.3 * (x1 - x0), 0.5 * (y0 + y1), z1 + height)
    look = (0.5 * (x0 + x1), 0.5 * (y0 + y1), 0.5 * (z0 + z1))
    s = perspective_projection(0.86, fov, nx, ny, pos, look, (0.0, 1.0, 0.0), sub_pixel_ray_args=sub)
    return {k: np.asarray(s[k]) for k in ("ray_x", "ray_y", "ray_z", "ray_mu", "ray_phi")}, s


def interior_rays(w, n, seed):
    """Rays that start inside the domain with arbitrary directions (up and down)."""
    rng = np.random.default_rng(seed)
    nx, ny, _ = w.shape
    x = rng.uniform(0, (nx - 1) * w.delx, n)
    y = rng.uniform(0, (ny - 1) * w.dely, n)
    z = rng.uniform(w.z[0], w.z[-1], n)
    mu = rng.uniform(-0.98, 0.98, n)
    mu[np.abs(mu) < 0.05] = 0.3
    phi = rng.uniform(0, 2 * np.pi, n)
    return {"ray_x": x, "ray_y": y, "ray_z": z, "ray_mu": mu, "ray_phi": phi}
