"""Small seeded scenes shared by the parity tests."""
import numpy as np

from at3d_amd import synthetic as S


def blob8(bc="periodic", nleg=64, seed=0):
    return S.lognormal_blob(n=8, radius=0.15, seed=seed, nleg=nleg, bc=bc)


def cloud_small(seed=1, n=12, nz=10):
    """Two species (cloud + Rayleigh) on an open grid, NMU 8 / NPHI 16."""
    w = S.cloud_field(nx=n, ny=n, nz=nz, dx=0.05, dz=0.1, z0=0.4, seed=seed, ext_scale=20.0,
                      nmu=8, nphi=16, solacc=1e-5)
    return w


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
