"""The allocation-free render entry point (at3d_render_noalloc) against
at3d_render, through the C ABI: bit-identical radiances on two calls in a
row, the second with a smaller input, and a request beyond the scratch size
refused before any device work."""
import ctypes as C

import numpy as np
import pytest

from at3d_amd import _lib
from at3d_amd import synthetic as S
from tests.scenes import cloud_l15, interior_rays, ortho_rays

pytestmark = [pytest.mark.gpu, pytest.mark.filterwarnings("ignore:RTE is not solved")]


def _render(entry, rte, rays):
    x = _lib.as_f32(rays["ray_x"])
    y = _lib.as_f32(rays["ray_y"])
    z = _lib.as_f32(rays["ray_z"])
    mu = np.ascontiguousarray(rays["ray_mu"], np.float64)
    phi = np.ascontiguousarray(rays["ray_phi"], np.float64)
    out = np.full(x.size, np.nan, np.float32)
    err = _lib.errbuf()
    rc = getattr(_lib.lib(), entry)(rte._state, x.size, x.ctypes.data, y.ctypes.data, z.ctypes.data,
                                    mu.ctypes.data, phi.ctypes.data, 0, 0, out.ctypes.data, err,
                                    len(err))
    return rc, out, err.value.decode()


@pytest.fixture(scope="module")
def solved():
    w = cloud_l15()
    rte = S.make_rte(w)
    rte.solve(3, verbose=False)
    return w, rte


def test_noalloc_render_matches_render_on_two_calls(solved):
    w, rte = solved
    assert _lib.lib().at3d_render_scratch_bytes() == 64 << 20
    big = interior_rays(w, 3000, seed=11)                      # 3000 directions
    small = ortho_rays(w, [(0.0, 0.0), (45.0, 30.0)])          # 2 directions
    assert small["ray_x"].size < big["ray_x"].size
    for rays in (big, small):
        rc, want, msg = _render("at3d_render", rte, rays)
        assert rc == 0, msg
        rc, got, msg = _render("at3d_render_noalloc", rte, rays)
        assert rc == 0, msg
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, want)


def test_noalloc_render_refuses_beyond_scratch_then_still_works(solved):
    w, rte = solved
    huge = interior_rays(w, 6000, seed=12)     # 6000 directions x every grid point > 64 MiB
    rc, _, msg = _render("at3d_render_noalloc", rte, huge)
    assert rc == 1   # AT3D_ERR_ARG
    assert "device scratch" in msg
    rays = ortho_rays(w, [(26.0, 95.0)])
    rc, want, msg = _render("at3d_render", rte, rays)
    assert rc == 0, msg
    rc, got, msg = _render("at3d_render_noalloc", rte, rays)
    assert rc == 0, msg
    assert np.array_equal(got, want)
