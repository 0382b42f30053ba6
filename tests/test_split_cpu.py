"""Adaptive cell splitting on the host against the reference's SPLIT_GRID
(shdomsub1.f:4703-5794), on identical inputs, bit for bit.

The oracle iterates a scene without splitting until SOLCRIT <= 0.1 (where
SOLUTION_ITERATIONS starts to split), then calls SPLIT_GRID one or more times
with decreasing splitting accuracies; the library's host splitting
(split_grid.cpp, reached through a small test hook linked against
libat3d_hip.so) runs on the same source and radiance.  Compared: the refined
grid (positions, cell corners, neighbour and tree pointers, flags), the
optics, source and radiance of the new points, SPLITCRIT, the out-of-memory
stop, and the refined grid's sweep order (SWEEPING_ORDER as the oracle's next
path integration leaves it in WORK1).
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from at3d_amd import synthetic as S
from tests.scenes import blob8, cloud_small, oo_blob, pol_blob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "at3d_amd")


@pytest.fixture(scope="module")
def hook(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(os.path.join(LIBDIR, "libat3d_hip.so")) or not os.path.exists(hipcc):
        pytest.skip("needs the built library and hipcc")
    out = str(tmp_path_factory.mktemp("hook") / "split_hook.so")
    src = os.path.join(ROOT, "tests", "native", "split_hook.cpp")
    cmd = [hipcc, "-O2", "-fPIC", "-shared", "-std=c++17", "--offload-arch=gfx950", "-x", "hip", src,
           "-o", out, f"-L{LIBDIR}", "-lat3d_hip", f"-Wl,-rpath,{LIBDIR}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return C.CDLL(out)


def _uniform_slab(n=6, nz=8):
    """Extinction varying in z only on a periodic grid: every cell of a layer
    has the same splitting criterion, so the order among equal criteria (the
    sort's) decides the point numbering."""
    z = (np.arange(nz) * 0.04).astype(np.float32)
    prof = np.array([0, 0, 5, 20, 40, 10, 2, 0], np.float32)[:nz]
    ext = np.broadcast_to(prof, (n, n, nz)).astype(np.float32).copy()
    alb = np.full(ext.shape, 0.99, np.float32)
    return S.Workload("slab", 0.05, 0.05, z, [(ext, alb, S.hg_legendre(0.85, 64))], bc="periodic",
                      nmu=8, nphi=16, solacc=1e-5)


def _dense(src, ptr, n, ld, k=0):
    """Stokes component k of a ragged (nstokes, maxiv) array as dense rows."""
    cnt = np.diff(np.asarray(ptr[:n + 1])).astype(np.int32)
    flat = np.asarray(src)[k]
    out = np.zeros((n, ld), np.float32)
    for p in range(n):
        out[p, :cnt[p]] = flat[ptr[p]:ptr[p] + cnt[p]]
    return out, cnt


CASES = {
    "blob8_open_to_out_of_memory": (lambda: blob8("open"), 0.03, [0.3, 0.1, 0.03]),
    "blob8_periodic": (lambda: blob8("periodic"), 0.03, [0.3, 0.05]),
    "cloud12_two_species": (cloud_small, 0.01, [0.2, 0.05, 0.01]),
    "uniform_slab_ties": (_uniform_slab, 0.01, [0.05, 0.01]),
    "polarized_blob8_open": (lambda: pol_blob(), 0.03, [0.3, 0.1, 0.03]),
    "polarized_blob8_periodic_oo": (lambda: oo_blob("periodic", nstokes=3), 0.03, [0.3, 0.05]),
    "interp_oo_open": (lambda: oo_blob("open"), 0.03, [0.3, 0.1, 0.03]),
}


@pytest.mark.parametrize("name", list(CASES))
def test_host_splitting_matches_reference(ref, hook, name):
    make, acc, accs = CASES[name]
    w = make()
    rte = S.make_rte(w, split_accuracy=acc)
    desc = rte._desc
    st = ref.setup(ref.scene_from_workload(w, splitacc=acc))
    ref.init_solution(st)
    assert (st.maxig, st.maxic, st.maxiv, st.maxido) == (desc.maxig, desc.maxic, desc.maxiv, desc.maxido)
    st.sc.splitacc = 0.0          # iterate on the base grid up to where splitting starts
    while st.solcrit > 0.1:
        ref.solution_iterations(st, st.iters + 1)
    d = st.d
    n0, nlm = d["npts"], d["nlm"]
    ld = (nlm + 63) // 64 * 64
    pol = d["nstokes"] == 3
    S0, ns0 = _dense(d["source"], d["shptr"], n0, ld)
    R0, nr0 = _dense(d["radiance"], d["rshptr"], n0, ld)
    SQU0 = RQU0 = None
    if pol:
        SQU0 = np.stack([_dense(d["source"], d["shptr"], n0, ld, k)[0] for k in (1, 2)])
        RQU0 = np.stack([_dense(d["radiance"], d["rshptr"], n0, ld, k)[0] for k in (1, 2)])
    crit_ref, oom = [], False
    for a in accs:
        c, oom = ref.split_grid(st, True, a, oom)
        crit_ref.append(c)
    N, NC = d["npts"], d["ncells"]
    assert N > n0

    cap_p, cap_c, npart = d["maxig"], d["maxic"], rte.plan_size("npart")
    sizes = np.zeros(3, np.int32)
    gp = np.zeros((cap_p, 3), np.float32)
    gptr = np.zeros((cap_c, 8), np.int32)
    nptr = np.zeros((cap_c, 6), np.int32)
    tptr = np.zeros((cap_c, 2), np.int32)
    cf = np.zeros(cap_c, np.int8)
    te = np.zeros(cap_p, np.float32)
    ex = np.zeros(npart * cap_p, np.float32)
    al = np.zeros(npart * cap_p, np.float32)
    So, Ro = np.zeros((cap_p, ld), np.float32), np.zeros((cap_p, ld), np.float32)
    nso, nro = np.zeros(cap_p, np.int32), np.zeros(cap_p, np.int32)
    crit = np.zeros(len(accs), np.float32)
    swo = np.zeros(8 * cap_p, np.int32)
    A = np.asarray(accs, np.float32)
    err = C.create_string_buffer(512)
    ptr = lambda a: C.c_void_p(a.ctypes.data)
    SQUo = np.zeros((2, cap_p, ld), np.float32)
    RQUo = np.zeros((2, cap_p, ld), np.float32)
    opt = lambda a: None if a is None else ptr(a)
    rc = hook.split_hook_pol(C.byref(desc), C.c_int(ld), ptr(S0), ptr(ns0), ptr(R0), ptr(nr0), opt(SQU0), opt(RQU0),
                             C.c_int(len(accs)), ptr(A), C.c_int(0), C.c_int64(cap_p), C.c_int64(cap_c), ptr(sizes),
                             ptr(gp), ptr(gptr), ptr(nptr), ptr(tptr), ptr(cf), ptr(te), ptr(ex), ptr(al), ptr(So),
                             ptr(nso), ptr(Ro), ptr(nro), ptr(SQUo) if pol else None, ptr(RQUo) if pol else None,
                             ptr(crit), ptr(swo), err, C.c_size_t(512))
    assert rc == 0, err.value
    assert (sizes[0], sizes[1], bool(sizes[2])) == (N, NC, oom)
    assert np.array_equal(crit, np.float32(crit_ref))
    # the library's 0-based pointers are the reference's 1-based ones minus 1
    # (-1 for none; -(c+2) for the reference's negative pointer -(c+1))
    assert np.array_equal(gp[:N], d["gridpos"][:, :N].T)
    assert np.array_equal(gptr[:NC] + 1, d["gridptr"][:, :NC].T)
    assert np.array_equal(nptr[:NC] + 1, d["neighptr"][:, :NC].T)
    assert np.array_equal(tptr[:NC] + 1, d["treeptr"][:, :NC].T)
    assert np.array_equal(cf[:NC].astype(np.int16), d["cellflags"][:NC])
    assert np.array_equal(te[:N], d["total_ext"][:N])
    assert np.array_equal(ex[:npart * N].reshape(npart, N), d["extinct"][:N, :].T)
    assert np.array_equal(al[:npart * N].reshape(npart, N), d["albedo"][:N, :].T)
    S1, ns1 = _dense(d["source"], d["shptr"], N, ld)
    R1, nr1 = _dense(d["radiance"], d["rshptr"], N, ld)
    assert np.array_equal(nso[:N], ns1) and np.array_equal(nro[:N], nr1)
    assert np.array_equal(So[:N], S1) and np.array_equal(Ro[:N], R1)
    if pol:
        for k in (1, 2):
            Sk, _ = _dense(d["source"], d["shptr"], N, ld, k)
            Rk, _ = _dense(d["radiance"], d["rshptr"], N, ld, k)
            assert np.array_equal(SQUo[k - 1, :N], Sk), k
            assert np.array_equal(RQUo[k - 1, :N], Rk), k
    # the refined grid's sweep order
    ref.solution_iterations(st, st.iters + 1)
    assert st.npts == N
    sw = st.work1[:8 * N].reshape(8, N)
    mine = swo[:8 * N].reshape(8, N)
    for j, ioct in enumerate([1, 5, 2, 6, 3, 7, 4, 8]):
        assert np.array_equal(mine[ioct - 1], sw[j] - 8), ioct
