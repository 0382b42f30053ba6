"""SHDOM property files: the four formats of READ_PROPERTY_SIZE /
READ_PROPERTIES (the standard one, extinction only 'E', tabulated 'T' and
polarized tabulated 'P') (src/polarized/shdomsub3.f:286-366,
:367-560), which the reference reaches through ``at3d.core`` (the verification
of tests/test_shdom.py:62-170 reads tests/data/rico32x36x26w672.prp this way).

Layout of both formats (Fortran list-directed records):
  line 1        the format letter ('T' or 'P'; anything else is not read here)
  NPX NPY NPZ
  DELX DELY ZLEVELS(1..NPZ)
  NUMPHASE
  per phase function: 'T': NUML CHI(1..NUML)
                      'P': six records JS NUML CHI(0..NUML), elements
                           a1 a2 a3 a4 b1 b2 in that order
  per property point: IX IY IZ TEMP EXT ALB IPHASE (1-based; points outside
                      the grid are skipped)
A list-directed READ starts at the next record and takes as many values as it
needs across records; what is left of its last record is skipped.

The Legendre coefficients include the (2l+1) factor, as the Scatterer tables
do.  NLEG grows to the largest NUML in the file (the reference takes its input
NLEG as a minimum, capped at MAXLEG, which READ_PROPERTY_SIZE makes the
largest NUML); terms past a table's NUML are zero, and a 'T' table's l = 0
term is 1.
"""
from __future__ import annotations

import gzip
from dataclasses import dataclass

import numpy as np

from .medium import Scatterer


class _Records:
    """Fortran list-directed reads over the lines of a file."""

    def __init__(self, path):
        opener = gzip.open if str(path).endswith(".gz") else open
        with opener(path, "rt") as f:
            self._lines = f.read().splitlines()
        self._i = 0

    def first_char(self) -> str:
        line = self._lines[self._i]
        self._i += 1
        return line[:1]

    def read(self, n: int) -> list[str] | None:
        """The next n values from the next record on (None at the end of
        the file before any value)."""
        out: list[str] = []
        while len(out) < n:
            if self._i >= len(self._lines):
                if not out:
                    return None
                raise ValueError("property file ends inside a record")
            out.extend(self._lines[self._i].replace(",", " ").split())
            self._i += 1
        return out[:n]


@dataclass
class PropertyFile:
    """The arrays READ_PROPERTIES returns (0-based numpy layout)."""
    proptype: str
    npx: int
    npy: int
    npz: int
    delx: float
    dely: float
    zlevels: np.ndarray      # (npz,) float32
    tempp: np.ndarray        # (npx, npy, npz) float32
    extinctp: np.ndarray     # (npx, npy, npz) float32
    albedop: np.ndarray      # (npx, npy, npz) float32
    iphasep: np.ndarray      # (npx, npy, npz) int32, 1-based
    legenp: np.ndarray       # (6, nleg+1, numphase) 'P' or (nleg+1, numphase) 'T', float32
    numphase: int
    nleg: int
    maxasym: float

    def scatterer(self) -> Scatterer:
        """The file's medium as one species (its table_index the file's
        phase function pointers)."""
        return Scatterer(self.extinctp, self.albedop, self.legenp, self.delx, self.dely, self.zlevels,
                         table_index=self.iphasep)


def read_property_file(path: str, nleg: int = 0) -> PropertyFile:
    """READ_PROPERTY_SIZE + READ_PROPERTIES for the 'T' and 'P' formats
    (a path ending in .gz is read through gzip).  ``nleg``: the minimum
    number of Legendre terms kept (the solver's NLEG)."""
    r = _Records(path)
    proptype = r.first_char()
    if proptype not in ("T", "P", "E"):
        r._i = 0            # the standard format has no format line
        return _read_standard(r, nleg)
    npx, npy, npz = (int(v) for v in r.read(3))
    hdr = r.read(2 + npz)
    delx, dely = float(np.float32(hdr[0])), float(np.float32(hdr[1]))
    zlevels = np.array(hdr[2:], np.float32)
    if proptype == "E":
        return _read_extinction_only(r, npx, npy, npz, delx, dely, zlevels, nleg)
    numphase = int(r.read(1)[0])
    if numphase < 1:
        raise ValueError("property file has no phase functions")
    tables = []
    for _ in range(numphase):
        if proptype == "T":
            numl = int(r.read(1)[0])
            r._i -= 1   # the record holds NUML and the coefficients
            vals = r.read(1 + numl)[1:]
            tables.append([np.array(vals, np.float32)])
        else:
            elems = []
            for j in range(6):
                head = r.read(2)
                js, numl = int(head[0]), int(head[1])
                if js != j + 1:
                    raise ValueError("polarized phase function element {} where {} was expected".format(js, j + 1))
                r._i -= 1
                elems.append(np.array(r.read(3 + numl)[2:], np.float32))
            tables.append(elems)
    numl_max = max(len(e) - (1 if proptype == "P" else 0) for t in tables for e in t)
    nl = max(int(nleg), numl_max)
    if proptype == "T":
        legenp = np.zeros((nl + 1, numphase), np.float32)
        for i, (e,) in enumerate(tables):
            legenp[0, i] = 1.0
            legenp[1:1 + len(e), i] = e
        maxasym = float(np.max(legenp[1] / np.float32(3.0))) if nl >= 1 else -1.0
    else:
        legenp = np.zeros((6, nl + 1, numphase), np.float32)
        for i, elems in enumerate(tables):
            for j, e in enumerate(elems):
                legenp[j, :len(e), i] = e
        maxasym = float(np.max(legenp[0, 1] / np.float32(3.0))) if nl >= 1 else -1.0
    shape = (npx, npy, npz)
    tempp = np.zeros(shape, np.float32)
    extinctp = np.zeros(shape, np.float32)
    albedop = np.zeros(shape, np.float32)
    iphasep = np.zeros(shape, np.int32)
    while True:
        rec = r.read(7)
        if rec is None:
            break
        ix, iy, iz = int(rec[0]), int(rec[1]), int(rec[2])
        if 1 <= ix <= npx and 1 <= iy <= npy and 1 <= iz <= npz:
            tempp[ix - 1, iy - 1, iz - 1] = np.float32(rec[3])
            extinctp[ix - 1, iy - 1, iz - 1] = np.float32(rec[4])
            albedop[ix - 1, iy - 1, iz - 1] = np.float32(rec[5])
            iphasep[ix - 1, iy - 1, iz - 1] = int(rec[6])
    if np.any(iphasep < 1) or np.any(iphasep > numphase):
        raise ValueError("property file leaves grid points without a valid phase function")
    return PropertyFile(proptype, npx, npy, npz, delx, dely, zlevels, tempp, extinctp, albedop, iphasep,
                        legenp, numphase, nl, maxasym)


def _read_extinction_only(r, npx, npy, npz, delx, dely, zlevels, nleg):
    """Format 'E': one temperature per level, then ALB NUML CHI(1..NUML) for
    the one phase function, then IX IY IZ EXT per point (IX IZ EXT when
    NPY = 1).  NLEG grows to NUML; as READ_PROPERTIES does
    (shdomsub3.f:430-465), coefficients past NUML are zeroed from index NUML
    on, so with NUML < NLEG the last coefficient read is zeroed too."""
    tz = np.array(r.read(npz), np.float32)
    head = r.read(2)
    alb, numl = np.float32(head[0]), int(head[1])
    r._i -= 1
    chi = np.array(r.read(2 + numl)[2:], np.float32)
    nl = max(int(nleg), numl)
    leg = np.zeros(nl + 1, np.float32)
    leg[0] = 1.0
    leg[1:1 + min(numl, nl)] = chi[:min(numl, nl)]
    for l in range(numl + 1, nl + 1):      # LEGENP(L) = 0 for L = NUML+1 .. NLEG (1-based)
        leg[l - 1] = 0.0
    maxasym = float(leg[1] / np.float32(3.0)) if nl >= 1 else -1.0
    shape = (npx, npy, npz)
    tempp = np.zeros(shape, np.float32)
    extinctp = np.zeros(shape, np.float32)
    albedop = np.zeros(shape, np.float32)
    iphasep = np.zeros(shape, np.int32)
    while True:
        rec = r.read(3 if npy == 1 else 4)
        if rec is None:
            break
        if npy == 1:
            ix, iy, iz, ext = int(rec[0]), 1, int(rec[1]), rec[2]
        else:
            ix, iy, iz, ext = int(rec[0]), int(rec[1]), int(rec[2]), rec[3]
        if 1 <= ix <= npx and 1 <= iy <= npy and 1 <= iz <= npz:
            tempp[ix - 1, iy - 1, iz - 1] = tz[iz - 1]
            extinctp[ix - 1, iy - 1, iz - 1] = np.float32(ext)
            albedop[ix - 1, iy - 1, iz - 1] = alb
            iphasep[ix - 1, iy - 1, iz - 1] = 1
    if np.any(iphasep < 1):
        raise ValueError("property file leaves grid points without properties")
    return PropertyFile("E", npx, npy, npz, delx, dely, zlevels, tempp, extinctp, albedop, iphasep,
                        leg[:, None], 1, nl, maxasym)


def _read_standard(r, nleg):
    """The standard format: IX IY IZ TEMP EXT ALB NUML CHI(1..NUML) per point,
    each point its own phase function (READ_PROPERTIES' NUMPHASE = 0; here
    one table per grid point, table_index = the point's index + 1, in the
    reference's point order IZ + NPZ (IY-1) + NPZ NPY (IX-1)).  The rows are
    the ones the reference's reader leaves: CHI(L) as term L - 1, term 0
    then set to 1 (so CHI(1) is dropped)."""
    npx, npy, npz = (int(v) for v in r.read(3))
    hdr = r.read(2 + npz)
    delx, dely = float(np.float32(hdr[0])), float(np.float32(hdr[1]))
    zlevels = np.array(hdr[2:], np.float32)
    recs = []
    while True:
        start = r._i
        head = r.read(7)
        if head is None:
            break
        numl = int(head[6])
        r._i = start
        recs.append(r.read(7 + numl))
    nl = max([int(nleg)] + [int(x[6]) for x in recs])
    shape = (npx, npy, npz)
    n = npx * npy * npz
    tempp = np.zeros(shape, np.float32)
    extinctp = np.zeros(shape, np.float32)
    albedop = np.zeros(shape, np.float32)
    iphasep = np.zeros(shape, np.int32)
    legenp = np.zeros((nl + 1, n), np.float32)
    maxasym = -1.0
    for x in recs:
        ix, iy, iz = int(x[0]), int(x[1]), int(x[2])
        if not (1 <= ix <= npx and 1 <= iy <= npy and 1 <= iz <= npz):
            continue
        k = iz + npz * (iy - 1) + npz * npy * (ix - 1) - 1
        tempp[ix - 1, iy - 1, iz - 1] = np.float32(x[3])
        extinctp[ix - 1, iy - 1, iz - 1] = np.float32(x[4])
        albedop[ix - 1, iy - 1, iz - 1] = np.float32(x[5])
        iphasep[ix - 1, iy - 1, iz - 1] = k + 1
        numl = int(x[6])
        chi = np.array(x[7:7 + min(numl, nl)], np.float32)
        # READ_PROPERTIES stores CHI(L) at LEGENP(L + KL), whose slot L + KL
        # is term L - 1 of the point's row, then sets term 0 to 1 and zeroes
        # LEGENP(L + KL) for L = NUML+1 .. NLEG (shdomsub3.f:536-548); the
        # same rows here, so a file reads as the reference reads it
        legenp[:, k] = 0.0
        legenp[:chi.size, k] = chi
        for l in range(numl + 1, nl + 1):
            legenp[l - 1, k] = 0.0
        maxasym = max(maxasym, float(legenp[1, k] / np.float32(3.0)))
        legenp[0, k] = 1.0
    if np.any(iphasep < 1):
        raise ValueError("property file leaves grid points without properties")
    return PropertyFile("S", npx, npy, npz, delx, dely, zlevels, tempp, extinctp, albedop, iphasep,
                        legenp, n, nl, maxasym)
