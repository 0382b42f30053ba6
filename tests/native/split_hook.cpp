// Test hook (not part of the library): runs the library's host-side adaptive
// cell splitting on a plan built from a scene descriptor, on solution fields
// given by the caller, so the CPU tests can compare it with the oracle's
// SPLIT_GRID on identical inputs.  Links against libat3d_hip.so.
#include <string.h>

#include <string>
#include <vector>

#include "../../at3d_amd/csrc/split_grid.h"

namespace {
template <typename T>
bool put(const std::vector<T>& v, T* dst, int64_t cap) {
  if ((int64_t)v.size() > cap) return false;
  if (!v.empty()) memcpy(dst, v.data(), v.size() * sizeof(T));
  return true;
}
}  // namespace

// Polarized fields (NSTOKES 3): SQU / RQU hold the Q then the U component,
// each [npts][ld] in and [cap_pts][ld] out; null for a scalar scene.
extern "C" int split_hook_pol(const at3d_scene_desc* desc, int ld, const float* S, const int32_t* ns, const float* R,
                              const int32_t* nr, const float* SQU, const float* RQU, int ncall,
                              const float* cursplitacc, int outofmem_in, int64_t cap_pts, int64_t cap_cells,
                              int32_t* sizes, float* gridpos, int32_t* gridptr, int32_t* neighptr, int32_t* treeptr,
                              int8_t* cellflags, float* total_ext, float* extinct, float* albedo, float* S_out,
                              int32_t* ns_out, float* R_out, int32_t* nr_out, float* SQU_out, float* RQU_out,
                              float* splitcrit, int32_t* sweepord, char* err, size_t errlen) {
  std::string msg;
  at3d::Plan p;
  int rc = at3d::build_plan(*desc, p, msg);
  if (!rc) rc = at3d::ensure_host_medium(p, msg);   // the splitting reads the host optics
  at3d::SplitFields f;
  f.ld = ld;
  if (!rc) {
    const size_t n = p.npts;
    f.S.assign(S, S + n * ld);
    f.R.assign(R, R + n * ld);
    f.DS.assign(n * ld, 0.f);
    f.ns.assign(ns, ns + n);
    f.nr.assign(nr, nr + n);
    f.nos.assign(n, 0);
    if (SQU) {
      f.nst = 3;
      for (int k = 0; k < 2; k++) {
        f.Squ[k].assign(SQU + k * n * ld, SQU + (k + 1) * n * ld);
        f.Rqu[k].assign(RQU + k * n * ld, RQU + (k + 1) * n * ld);
        f.DSqu[k].assign(n * ld, 0.f);
      }
    }
  }
  bool outofmem = outofmem_in != 0;
  for (int k = 0; k < ncall && !rc; k++) rc = at3d::split_grid(p, f, true, outofmem, cursplitacc[k], splitcrit[k], msg);
  if (!rc) {
    sizes[0] = p.npts;
    sizes[1] = p.ncells;
    sizes[2] = outofmem ? 1 : 0;
    bool ok = put(p.gridpos, gridpos, cap_pts * 3) && put(p.gridptr, gridptr, cap_cells * 8) &&
              put(p.neighptr, neighptr, cap_cells * 6) && put(p.treeptr, treeptr, cap_cells * 2) &&
              put(p.cellflags, cellflags, cap_cells) && put(p.total_ext, total_ext, cap_pts) &&
              put(p.extinct, extinct, cap_pts * p.npart) && put(p.albedo, albedo, cap_pts * p.npart) &&
              put(f.S, S_out, cap_pts * ld) && put(f.ns, ns_out, cap_pts) && put(f.R, R_out, cap_pts * ld) &&
              put(f.nr, nr_out, cap_pts);
    for (int k = 0; ok && SQU && k < 2; k++)
      ok = put(f.Squ[k], SQU_out + k * cap_pts * ld, cap_pts * ld) &&
           put(f.Rqu[k], RQU_out + k * cap_pts * ld, cap_pts * ld);
    if (!ok) {
      rc = 1;
      msg = "output capacity too small";
    }
  }
  // the refined grid's sweep order, per octant (IOCT - 1)
  if (!rc && p.npts != (int)(f.ns.size()) ) {
    rc = 1;
    msg = "fields and grid disagree";
  }
  if (!rc) rc = at3d::plan_refresh_points(p, msg);
  if (!rc && sweepord)
    for (int o = 0; o < 8; o++)
      memcpy(sweepord + (size_t)o * p.npts, p.oct[o].sweepord.data(), (size_t)p.npts * sizeof(int32_t));
  if (rc && err && errlen) {
    strncpy(err, msg.c_str(), errlen - 1);
    err[errlen - 1] = 0;
  }
  return rc;
}

extern "C" int split_hook(const at3d_scene_desc* desc, int ld, const float* S, const int32_t* ns, const float* R,
                          const int32_t* nr, int ncall, const float* cursplitacc, int outofmem_in,
                          int64_t cap_pts, int64_t cap_cells, int32_t* sizes, float* gridpos, int32_t* gridptr,
                          int32_t* neighptr, int32_t* treeptr, int8_t* cellflags, float* total_ext,
                          float* extinct, float* albedo, float* S_out, int32_t* ns_out, float* R_out,
                          int32_t* nr_out, float* splitcrit, int32_t* sweepord, char* err, size_t errlen) {
  return split_hook_pol(desc, ld, S, ns, R, nr, nullptr, nullptr, ncall, cursplitacc, outofmem_in, cap_pts, cap_cells,
                        sizes, gridpos, gridptr, neighptr, treeptr, cellflags, total_ext, extinct, albedo, S_out,
                        ns_out, R_out, nr_out, nullptr, nullptr, splitcrit, sweepord, err, errlen);
}
