"""Saved-solution layout and load checks that need no GPU: the grid part of
RTE.save_solution against the reference's own grid arrays, and the errors
load_solution raises for an inconsistent dataset (at3d/solver.py:1552-1605)."""
import numpy as np
import pytest

from at3d_amd import synthetic as S
from tests.scenes import blob8


def test_saved_grid_matches_reference_layout(ref):
    w = blob8("open")
    rte = S.make_rte(w)
    ds = rte.save_solution(save_radiances=False)
    st = ref.setup(ref.scene_from_workload(w))
    assert (ds["npts"], ds["ncells"]) == (st.npts, st.ncells)
    assert (ds["nx"], ds["ny"], ds["nz"]) == (st.nx, st.ny, st.nz)
    assert (ds["ml"], ds["mm"]) == (st.ml, st.mm)
    assert np.array_equal(ds["xgrid"], st.xgrid) and np.array_equal(ds["ygrid"], st.ygrid)
    assert np.array_equal(ds["zgrid"], st.zgrid)
    assert np.array_equal(ds["gridpos"], st.gridpos[:, :st.npts])
    assert np.array_equal(ds["gridptr"], st.gridptr[:, :st.ncells])
    assert np.array_equal(ds["neighptr"], st.neighptr[:, :st.ncells])
    assert np.array_equal(ds["cellflags"], st.cellflags[:st.ncells])
    assert "source" not in ds


def _radiance_vars(rte):
    n = rte.plan_size("npts")
    return {"shptr": np.arange(n + 1, dtype=np.int32), "rshptr": np.arange(n + 2, dtype=np.int32),
            "source": np.zeros((1, n), np.float32), "radiance": np.zeros((1, n + 1), np.float32),
            "fluxes": np.zeros((2, n), np.float32)}


def test_load_solution_rejects_inconsistent_datasets():
    a = S.make_rte(blob8("open"))
    ds = a.save_solution(save_radiances=False)
    ds.update(_radiance_vars(a))
    other_grid = S.make_rte(S.lognormal_blob(n=9, radius=0.15, seed=0, nleg=64, bc="open", nmu=8,
                                             nphi=16))
    with pytest.raises(ValueError, match="Incompatible grid sizes"):
        other_grid.load_solution(ds)
    finer_angles = S.make_rte(blob8("open", nmu=16, nphi=32))
    coarse = dict(ds)
    coarse["ml"], coarse["mm"] = 3, 3
    with pytest.raises(ValueError, match="ml, mm"):
        finer_angles.load_solution(coarse)
    bigger = dict(ds)
    bigger["ml"] = 40
    with pytest.raises(ValueError, match="Incompatible SphericalHarmonics"):
        a.load_solution(bigger)
    moved = dict(ds)
    moved["zgrid"] = ds["zgrid"] + np.float32(0.01)
    with pytest.raises(ValueError, match="Incompatible grids"):
        a.load_solution(moved)
    missing = {k: v for k, v in ds.items() if k != "rshptr"}
    with pytest.raises(KeyError, match="rshptr"):
        a.load_solution(missing)
    a.load_solution(ds, load_radiance=False)    # the grid alone is consistent
    a.load_solution(ds)
