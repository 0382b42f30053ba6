"""Writes tests/golden/shdom_small.npz: outputs of the reference SHDOM (the
Fortran oracle built under oracle/_ref) on the small seeded scenes of
tests/scenes.py.  Run from the repository root after `make -C oracle`:

    python -m tests.golden.make_golden

Stored per scene (prefix "<scene>/"): a checksum of the generated inputs, the
source after one iteration, the converged source, radiance and fluxes with the
iteration count and criterion, rendered radiances for orthographic views and
for interior rays in all directions (with the rays), and the Levis gradient
and L2 cost with respect to the first species' extinction and albedo for
measurements 1.1 x the view radiances.
"""
import hashlib
import os

import numpy as np

from oracle import ref_shdom as ref
from tests import scenes

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "shdom_small.npz")

SCENES = {"blob8_periodic": lambda: scenes.blob8("periodic"),
          "blob8_open": lambda: scenes.blob8("open"),
          "cloud12_two_species": scenes.cloud_small}
VIEWS = [(0.0, 0.0), (45.0, 30.0), (60.0, 200.0), (26.0, 95.0)]


def input_checksum(w):
    h = hashlib.sha256()
    for ext, alb, leg in w.species:
        for a in (ext, alb, leg):
            h.update(np.ascontiguousarray(a, np.float32).tobytes())
    h.update(np.ascontiguousarray(w.z, np.float32).tobytes())
    h.update(repr((w.delx, w.dely, w.bc, w.nmu, w.nphi, w.solacc, w.solarmu, w.solaraz_deg,
                   w.gndalbedo)).encode())
    return h.hexdigest()


def views(w):
    return scenes.ortho_rays(w, VIEWS)


def interior(w):
    return scenes.interior_rays(w, 150, seed=3)


def dense(st, which):
    if which == "source":
        return ref.dense_sh(st.shptr, st.source, st.npts, st.nlm)
    return ref.dense_sh(st.rshptr, st.radiance, st.npts, st.nlm)


def main():
    out = {}
    for name, make in SCENES.items():
        w = make()
        out[f"{name}/checksum"] = np.array(input_checksum(w))
        st = ref.setup(ref.scene_from_workload(w))
        ref.init_solution(st)
        ref.solution_iterations(st, 1)
        out[f"{name}/source_iter1"] = dense(st, "source")
        st = ref.setup(ref.scene_from_workload(w))
        ref.init_solution(st)
        ref.solution_iterations(st, 100)
        out[f"{name}/source"] = dense(st, "source")
        out[f"{name}/radiance"] = dense(st, "radiance")
        out[f"{name}/fluxes"] = np.ascontiguousarray(st.fluxes[:, :st.npts], np.float32)
        out[f"{name}/iters"] = np.array(st.iters)
        out[f"{name}/solcrit"] = np.array(st.solcrit, np.float32)
        for tag, rays in (("views", views(w)), ("interior", interior(w))):
            for k, v in rays.items():
                out[f"{name}/{tag}/{k}"] = np.asarray(v)
            out[f"{name}/{tag}/I"] = ref.render(st, rays["ray_x"], rays["ray_y"], rays["ray_z"],
                                                rays["ray_mu"], rays["ray_phi"])
        shape = w.shape
        nleg = w.species[0][2].shape[0]
        ref.prepare_derivatives(st, [
            ref.Derivative(0, np.ones(shape), np.zeros(shape), np.zeros((nleg, 1)), np.ones(shape, np.int32)),
            ref.Derivative(0, np.zeros(shape), np.ones(shape), np.zeros((nleg, 1)), np.ones(shape, np.int32))])
        v = views(w)
        meas = (out[f"{name}/views/I"] * 1.1).astype(np.float32)
        g, c, _ = ref.levis_gradient(st, v["ray_x"], v["ray_y"], v["ray_z"], v["ray_mu"], v["ray_phi"], meas)
        out[f"{name}/grad"] = g
        out[f"{name}/cost"] = np.array(c)
    np.savez_compressed(OUT, **out)
    print(OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
