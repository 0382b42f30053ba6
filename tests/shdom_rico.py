"""The reference's original-SHDOM verification case (tests/test_shdom.py:175-277):
the polarized RICO cloud of rico32x36x26w672.prp at 0.672 um, NMU 8 / NPHI 16,
adaptive splitting at 0.1, SH truncation at 0.01, solution accuracy 1e-4,
INTERPMETHOD 'OO', sun at mu0 = 0.5, Lambertian surface 0.05, and the radiances
SHDOM wrote at z = 1 km for 47 x 53 pixels in 5 directions (rendered without
CORRECTINTERPOLATE).  The expected values are SHDOM's own output files
(tests/golden/shdom_rico/, gzip copies)."""
import gzip
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "shdom_rico")
PRP = os.path.join(HERE, "rico32x36x26w672.prp.gz")

CONFIG = dict(num_mu_bins=8, num_phi_bins=16, split_accuracy=0.1, spherical_harmonics_accuracy=0.01,
              solution_accuracy=1e-4, tautol=0.1)
ITERATIONS = 18       # NUMBER_ITERATIONS of rico32x36x26w672ar.out
NPTS = 32809          # NPTS of the same header
# tolerances of the reference's test (test_shdom.py:262-275)
SOURCE_ATOL = 5e-7
STOKES_ATOL = (3e-3, 2e-4, 7e-5)


def parse(name, comment):
    """Rows of numbers of an SHDOM output file, lines holding `comment` skipped
    (the reference's parse_shdom_output)."""
    rows = []
    with gzip.open(os.path.join(HERE, name + ".gz"), "rt") as f:
        for line in f:
            if comment not in line:
                rows.append(np.array(line.split(), np.float64))
    return np.array(rows)


def source_truth():
    """(3, NPTS): the first NPTS entries of SHDOM's SOURCE(NSTOKES, *) array."""
    return parse("shdom_verification_source_out.out", "*").T


def radiance_truth():
    """(12455, 5): X, Y, I, Q, U per pixel, directions in the rays' order."""
    return parse("rico32x36x26w672ar.out", "!")


def rays():
    """The pixels of the radiance file: x, y, z, mu, phi (radians)."""
    x, y = np.meshgrid(np.arange(0.0, 0.62, 0.0133), np.arange(0.0, 0.70, 0.0133))
    x, y = x.ravel(), y.ravel()
    mu = np.array([1.0, 0.5, 0.2, 0.5, 0.2])
    phi = np.array([0.0, 0.0, 0.0, 90.0, 90.0])
    return (np.tile(x, 5), np.tile(y, 5), np.ones(5 * x.size), np.repeat(mu, x.size),
            np.deg2rad(np.repeat(phi, x.size)))


def flat_source(S, ns, n):
    """The first n entries of the ragged source of dense (3, npts, ld) rows
    with per-point counts ns, as the reference lays out SOURCE."""
    out = np.zeros((3, n), np.float32)
    pos, i = 0, 0
    while pos < n:
        k = min(int(ns[i]), n - pos)
        out[:, pos:pos + k] = S[:, i, :k]
        pos += k
        i += 1
    return out
