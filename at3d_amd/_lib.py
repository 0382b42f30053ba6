"""ctypes binding of libat3d_hip.so (declarations in include/at3d_hip.h).

The library is built in-tree (``python -m at3d_amd.build``); importing this
module never builds it and never falls back to anything else: a missing
library raises ``ImportError``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libat3d_hip.so")

_f32p = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)


class SceneDesc(C.Structure):
    _fields_ = [
        ("npx", C.c_int32), ("npy", C.c_int32), ("npz", C.c_int32),
        ("delx", C.c_float), ("dely", C.c_float),
        ("zlevels", _f32p),
        ("npart", C.c_int32),
        ("extinctp", _f32p), ("albedop", _f32p), ("iphasep", _i32p), ("phasewtp", _f32p),
        ("numphase", C.c_int32), ("nlegp", C.c_int32), ("legenp", _f32p),
        ("solarflux", C.c_float), ("solarmu", C.c_float), ("solaraz", C.c_float),
        ("skyrad", C.c_float), ("gndalbedo", C.c_float), ("wavelength", C.c_float),
        ("nmu", C.c_int32), ("nphi", C.c_int32), ("open_bc", C.c_int32), ("ip_flag", C.c_int32),
        ("deltam", C.c_int32), ("accelflag", C.c_int32), ("high_order_radiance", C.c_int32),
        ("iterfixsh", C.c_int32), ("angle_set", C.c_int32),
        ("shacc", C.c_float), ("solacc", C.c_float), ("splitacc", C.c_float),
        ("transmin", C.c_float), ("tautol", C.c_double), ("transcut", C.c_double),
        ("maxig", C.c_int32), ("maxic", C.c_int32), ("maxiv", C.c_int32), ("maxido", C.c_int32),
        ("device_medium", C.c_int32), ("nstokes", C.c_int32), ("nmicro", C.c_int32),
        ("interpmethod", C.c_int32),
    ]


class DerivDesc(C.Structure):
    _fields_ = [
        ("numder", C.c_int32), ("partder", _i32p), ("dext", _f32p), ("dalb", _f32p),
        ("doexact", _i32p), ("dnumphase", C.c_int32), ("dleg", _f32p), ("diphasep", _i32p),
        ("dphasewtp", _f32p),
    ]


class SHDOMError(RuntimeError):
    """Raised when the engine returns a nonzero status (mirrors at3d.exceptions.SHDOMError)."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(f"{LIBPATH} is not built: run `python -m at3d_amd.build`")
        # torch (where installed) first: the HIP runtime it bundles sees no
        # device once this library's runtime has initialised the GPU in the
        # same process, while this library works in either order
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIBPATH)
        vp = C.c_void_p
        L.at3d_version.restype = C.c_char_p
        L.at3d_device_trim.restype = C.c_int64
        L.at3d_plan_create.argtypes = [C.POINTER(SceneDesc), C.POINTER(vp), C.c_char_p, C.c_size_t]
        L.at3d_plan_destroy.argtypes = [vp]
        L.at3d_plan_size.argtypes = [vp, C.c_char_p]
        L.at3d_plan_size.restype = C.c_int64
        L.at3d_plan_get.argtypes = [vp, C.c_char_p, vp, C.c_size_t, C.c_char_p, C.c_size_t]
        L.at3d_state_create.argtypes = [vp, C.POINTER(vp), C.c_char_p, C.c_size_t]
        L.at3d_state_destroy.argtypes = [vp]
        L.at3d_init_solution.argtypes = [vp, C.c_char_p, C.c_size_t]
        L.at3d_init_solution_from.argtypes = [vp, vp, C.c_char_p, C.c_size_t]
        L.at3d_solve.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(C.c_float),
                                 C.POINTER(C.c_int32), C.c_char_p, C.c_size_t]
        L.at3d_state_get.argtypes = [vp, C.c_char_p, vp, C.c_size_t, C.c_char_p, C.c_size_t]
        L.at3d_state_set_option.argtypes = [vp, C.c_char_p, C.c_int32, C.c_char_p, C.c_size_t]
        L.at3d_render.argtypes = [vp, C.c_int64, vp, vp, vp, vp, vp, C.c_int32, C.c_int32, vp,
                                  C.c_char_p, C.c_size_t]
        L.at3d_state_load.argtypes = [vp, vp, vp, vp, vp, vp, C.c_char_p, C.c_size_t]
        L.at3d_state_save.argtypes = [vp, vp, vp, vp, C.c_int64, vp, C.c_int64, vp, C.c_char_p,
                                      C.c_size_t]
        L.at3d_state_sh_terms.argtypes = [vp, C.c_int32, vp, vp, C.c_char_p, C.c_size_t]
        L.at3d_render_noalloc.argtypes = [vp, C.c_int64, vp, vp, vp, vp, vp, C.c_int32, C.c_int32,
                                          vp, C.c_char_p, C.c_size_t]
        L.at3d_render_scratch_bytes.argtypes = []
        L.at3d_render_scratch_bytes.restype = C.c_int64
        L.at3d_levis_gradient.argtypes = [vp, C.POINTER(DerivDesc), C.c_int64, vp, vp, vp, vp, vp,
                                          C.c_int64, vp, vp, vp, vp, vp, C.c_char_p, C.c_int32,
                                          vp, vp, vp, C.c_char_p, C.c_size_t]
        L.at3d_levis_gradient_jacobian.argtypes = [vp, C.POINTER(DerivDesc), C.c_int64, vp, vp, vp, vp,
                                                   vp, C.c_int64, vp, vp, vp, vp, vp, C.c_char_p,
                                                   C.c_int32, C.c_int32, vp, vp, vp, vp, vp,
                                                   C.c_char_p, C.c_size_t]
        L.at3d_state_timing.argtypes = [vp, C.c_char_p]
        L.at3d_state_timing.restype = C.c_double
        L.at3d_state_size.argtypes = [vp, C.c_char_p]
        L.at3d_state_size.restype = C.c_int64
        L.at3d_state_set_grid.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, vp, vp, vp, C.c_char_p,
                                          C.c_size_t]
        L.at3d_version.argtypes = []
        L.at3d_device_trim.argtypes = []
        _lib = L
    return _lib


def check(rc: int, err) -> None:
    if rc != 0:
        raise SHDOMError(rc, err.value.decode(errors="replace"))


def errbuf():
    return C.create_string_buffer(1024)


# symbols every build must export (checked by the CPU test-suite)
EXPORTS = ["at3d_version", "at3d_plan_create", "at3d_plan_destroy", "at3d_plan_size",
           "at3d_plan_get", "at3d_state_create", "at3d_state_destroy", "at3d_init_solution",
           "at3d_solve", "at3d_state_get", "at3d_state_set_option", "at3d_state_timing",
           "at3d_render", "at3d_levis_gradient", "at3d_levis_gradient_jacobian",
           "at3d_init_solution_from", "at3d_device_trim", "at3d_state_load", "at3d_state_save",
           "at3d_state_sh_terms", "at3d_render_noalloc",
           "at3d_render_scratch_bytes", "at3d_state_size", "at3d_state_set_grid"]


def as_f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def as_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
