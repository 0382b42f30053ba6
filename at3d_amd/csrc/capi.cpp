// extern "C" entry points of libat3d_hip.so (declared in include/at3d_hip.h).
#include <hip/hip_runtime.h>
#include <string.h>

#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/at3d_hip.h"
#include "host_plan.h"
#include "ring_plan.h"
#include "devmem.h"
#include "solver_state.h"

struct at3d_plan {
  at3d::Plan p;
  // the host direct flux, made on first request (solves make their own)
  mutable std::mutex m;
  mutable std::vector<float> dirflux;
};
struct at3d_state {
  at3d::State s;
};

namespace {
void set_err(char* err, size_t errlen, const std::string& msg) {
  if (!err || errlen == 0) return;
  size_t n = msg.size() < errlen - 1 ? msg.size() : errlen - 1;
  memcpy(err, msg.data(), n);
  err[n] = 0;
}

template <typename T>
int copy_out(const std::vector<T>& v, void* dst, size_t nbytes, char* err, size_t errlen) {
  const size_t need = v.size() * sizeof(T);
  if (nbytes < need) {
    set_err(err, errlen, "destination buffer too small: need " + std::to_string(need) + " bytes");
    return AT3D_ERR_ARG;
  }
  if (need) memcpy(dst, v.data(), need);
  return AT3D_OK;
}
}  // namespace

extern "C" {

const char* at3d_version(void) { return "at3d_hip 0.1 (gfx950)"; }

int64_t at3d_device_trim(void) {
  const size_t held = at3d::dev_cached_bytes();
  at3d::dev_trim();
  return (int64_t)held;
}

int at3d_plan_create(const at3d_scene_desc* desc, at3d_plan** out, char* err, size_t errlen) {
  if (!desc || !out) {
    set_err(err, errlen, "null argument");
    return AT3D_ERR_ARG;
  }
  *out = nullptr;
  at3d_plan* h = new (std::nothrow) at3d_plan;
  if (!h) {
    set_err(err, errlen, "out of host memory");
    return AT3D_ERR_MEMORY;
  }
  std::string msg;
  int rc;
  try {
    rc = at3d::build_plan(*desc, h->p, msg);
  } catch (const std::bad_alloc&) {
    rc = AT3D_ERR_MEMORY;
    msg = "out of host memory building the plan";
  }
  if (rc) {
    set_err(err, errlen, msg);
    delete h;
    return rc;
  }
  *out = h;
  return AT3D_OK;
}

void at3d_plan_destroy(at3d_plan* plan) { delete plan; }

int64_t at3d_plan_size(const at3d_plan* h, const char* key) {
  if (!h || !key) return -1;
  const at3d::Plan& p = h->p;
  const std::string k(key);
  if (k == "npts") return p.npts;
  if (k == "ncells") return p.ncells;
  if (k == "nx") return p.nx;
  if (k == "ny") return p.ny;
  if (k == "nx1") return p.nx1;
  if (k == "ny1") return p.ny1;
  if (k == "nz") return p.nz;
  if (k == "nlm") return p.nlm;
  if (k == "ml") return p.ml;
  if (k == "mm") return p.mm;
  if (k == "nmu") return p.nmu;
  if (k == "nang") return p.nang;
  if (k == "nphi0max") return p.nphi0max;
  if (k == "bcflag") return p.bcflag;
  if (k == "ipflag") return p.ipflag;
  if (k == "ntop") return (int64_t)p.bctop.size();
  if (k == "nbot") return (int64_t)p.bcbot.size();
  if (k == "nlevels_max") return p.nlevels_max;
  if (k == "footprint_bytes") return (int64_t)at3d::plan_footprint_bytes(p);
  if (k == "npart") return p.npart;
  if (k == "numphase") return p.numphase;
  if (k == "nleg") return p.nleg;
  if (k == "nlegp") return p.nlegp;
  if (k == "nscatangle") return p.nscatangle;
  if (k == "nstokes") return p.nst;
  if (k == "nslow") {
    int64_t s = 0;
    for (int o = 0; o < 8; o++) s += p.oct[o].nslow;
    return s;
  }
  return -1;
}

int at3d_plan_get(const at3d_plan* h, const char* name, void* dst, size_t nbytes, char* err,
                  size_t errlen) {
  if (!h || !name || (!dst && nbytes)) {
    set_err(err, errlen, "null argument");
    return AT3D_ERR_ARG;
  }
  const at3d::Plan& p = h->p;
  const std::string k(name);
  // (the ring layout groups points by source class, made from the medium)
  if (k == "total_ext" || k == "extinct" || k == "albedo" || k == "radiance0" || k == "chi" || k == "alb" ||
      k.compare(0, 5, "ring_") == 0) {
    std::string msg;
    const int rc = at3d::ensure_host_medium(p, msg);
    if (rc) {
      set_err(err, errlen, msg);
      return rc;
    }
  }
  if (k == "gridpos") return copy_out(p.gridpos, dst, nbytes, err, errlen);
  if (k == "gridptr") return copy_out(p.gridptr, dst, nbytes, err, errlen);
  if (k == "neighptr") return copy_out(p.neighptr, dst, nbytes, err, errlen);
  if (k == "cellflags") return copy_out(p.cellflags, dst, nbytes, err, errlen);
  if (k == "xgrid") return copy_out(p.xgrid, dst, nbytes, err, errlen);
  if (k == "ygrid") return copy_out(p.ygrid, dst, nbytes, err, errlen);
  if (k == "zgrid") return copy_out(p.zgrid, dst, nbytes, err, errlen);
  if (k == "mu") return copy_out(p.mu, dst, nbytes, err, errlen);
  if (k == "wtmu") return copy_out(p.wtmu, dst, nbytes, err, errlen);
  if (k == "nphi0") return copy_out(p.nphi0, dst, nbytes, err, errlen);
  if (k == "phi") return copy_out(p.phi, dst, nbytes, err, errlen);
  if (k == "wtdo") return copy_out(p.wtdo, dst, nbytes, err, errlen);
  if (k == "ymat") return copy_out(p.ymat, dst, nbytes, err, errlen);
  if (k == "wmat") return copy_out(p.wmat, dst, nbytes, err, errlen);
  if (k == "ylmsun") return copy_out(p.ylmsun, dst, nbytes, err, errlen);
  // polarized plans: the six-element transform tables, the I-Q phase table,
  // the a2 / a3 / b1 source terms and the b1 single-scatter terms
  if (k == "ypol") return copy_out(p.ypol, dst, nbytes, err, errlen);
  if (k == "wpol") return copy_out(p.wpol, dst, nbytes, err, errlen);
  if (k == "phasetab") return copy_out(p.phasetab, dst, nbytes, err, errlen);
  if (k == "phasetab2") return copy_out(p.phasetab2, dst, nbytes, err, errlen);
  if (k == "chip") return copy_out(p.chip, dst, nbytes, err, errlen);
  if (k == "legen") return copy_out(p.legen, dst, nbytes, err, errlen);
  if (k == "legen2") return copy_out(p.legen2, dst, nbytes, err, errlen);
  if (k == "legen3") return copy_out(p.legen3, dst, nbytes, err, errlen);
  if (k == "legen5") return copy_out(p.legen5, dst, nbytes, err, errlen);
  if (k == "tms5") return copy_out(p.tms5, dst, nbytes, err, errlen);
  if (k == "total_ext") return copy_out(p.total_ext, dst, nbytes, err, errlen);
  if (k == "extinct") return copy_out(p.extinct, dst, nbytes, err, errlen);
  if (k == "albedo") return copy_out(p.albedo, dst, nbytes, err, errlen);
  if (k == "dirflux") {
    std::lock_guard<std::mutex> lock(h->m);
    if (h->dirflux.empty() && p.npts > 0) {
      std::string msg;
      const int rc = at3d::plan_direct_host(p, h->dirflux, msg);
      if (rc) {
        set_err(err, errlen, msg);
        return rc;
      }
    }
    return copy_out(h->dirflux, dst, nbytes, err, errlen);
  }
  if (k == "radiance0") return copy_out(p.radiance0, dst, nbytes, err, errlen);
  if (k == "chi") return copy_out(p.chi, dst, nbytes, err, errlen);
  if (k == "alb") return copy_out(p.alb, dst, nbytes, err, errlen);
  if (k == "bcptr_top") return copy_out(p.bctop, dst, nbytes, err, errlen);
  if (k == "sunpath_off" || k == "sunpath_ptr" || k == "sunpath_val") {
    // per RTE point, the property points on its path to the sun and the
    // derivative of the slant optical path with respect to their extinction
    std::vector<int32_t> off, ptr;
    std::vector<float> val;
    std::string msg;
    const int rc = at3d::plan_sun_paths(p, off, ptr, val, msg);
    if (rc) {
      set_err(err, errlen, msg);
      return rc;
    }
    if (k == "sunpath_off") return copy_out(off, dst, nbytes, err, errlen);
    if (k == "sunpath_ptr") return copy_out(ptr, dst, nbytes, err, errlen);
    return copy_out(val, dst, nbytes, err, errlen);
  }
  if (k == "bcptr_bot") return copy_out(p.bcbot, dst, nbytes, err, errlen);
  if (k == "ring_level" || k == "ring_pos" || k == "ring_staged" || k == "ring_sizes") {
    // the level-ordered sweep's layout ([8][npts] per octant, zeros for
    // octants without ordinates; ring_sizes = ring slots, points per sector,
    // levels of each octant)
    bool used[8] = {};
    for (const at3d::Ordinate& od : p.ords) used[od.g.ioct] = true;
    at3d::RingPlan rp;
    std::string msg;
    if (!at3d::build_ring_plan(p, used, at3d::GFX950_LDS_BYTES, rp, msg)) {
      set_err(err, errlen, msg);
      return AT3D_ERR_NUMERIC;
    }
    std::vector<int32_t> all;
    if (k == "ring_sizes") {
      all.push_back(rp.slots);
      all.push_back(rp.sector_points);
      for (int o = 0; o < 8; o++) all.push_back(rp.oct[o].nlev);
    } else {
      all.assign((size_t)8 * p.npts, 0);
      for (int o = 0; o < 8; o++) {
        if (!used[o]) continue;
        const at3d::RingOctant& ro = rp.oct[o];
        for (int i = 0; i < p.npts; i++)
          all[(size_t)o * p.npts + i] = k == "ring_level" ? ro.level[i] : k == "ring_pos" ? ro.pos[i]
                                                                                        : ro.staged[i];
      }
    }
    return copy_out(all, dst, nbytes, err, errlen);
  }
  if (k == "sweepord" || k == "level" || k == "lvl_pts") {
    // per octant, concatenated: [8][npts] for sweepord and level; lvl_pts is
    // each octant's non-preset points in (level, rank) order, ~point for the
    // multi-cell traces, octant after octant
    std::vector<int32_t> all;
    all.reserve((size_t)8 * p.npts);
    for (int o = 0; o < 8; o++) {
      const std::vector<int32_t>& v = (k == "sweepord") ? p.oct[o].sweepord
                                      : (k == "level")  ? p.oct[o].level
                                                        : p.oct[o].lvl_pts;
      all.insert(all.end(), v.begin(), v.end());
    }
    return copy_out(all, dst, nbytes, err, errlen);
  }
  set_err(err, errlen, "unknown plan array '" + k + "'");
  return AT3D_ERR_ARG;
}

int at3d_state_create(const at3d_plan* plan, at3d_state** out, char* err, size_t errlen) {
  if (!plan || !out) {
    set_err(err, errlen, "null argument");
    return AT3D_ERR_ARG;
  }
  *out = nullptr;
  int ndev = 0;
  const hipError_t dc = hipGetDeviceCount(&ndev);
  if (dc != hipSuccess || ndev < 1) {
    set_err(err, errlen, std::string("no HIP device available (hipGetDeviceCount: ") + hipGetErrorString(dc) +
                             ", " + std::to_string(ndev) + " devices)");
    return AT3D_ERR_DEVICE;
  }
  at3d_state* h = new (std::nothrow) at3d_state;
  if (!h) {
    set_err(err, errlen, "out of host memory");
    return AT3D_ERR_MEMORY;
  }
  std::string msg;
  int rc;
  try {
    rc = at3d::state_create(plan->p, h->s, msg);
  } catch (const std::bad_alloc&) {
    rc = AT3D_ERR_MEMORY;
    msg = "out of host memory creating the device state";
  }
  if (rc) {
    at3d::state_free(h->s);
    delete h;
    set_err(err, errlen, msg);
    return rc;
  }
  *out = h;
  return AT3D_OK;
}

void at3d_state_destroy(at3d_state* st) {
  if (!st) return;
  at3d::state_free(st->s);
  delete st;
}

int at3d_init_solution(at3d_state* st, char* err, size_t errlen) {
  if (!st) {
    set_err(err, errlen, "null state");
    return AT3D_ERR_ARG;
  }
  std::string msg;
  int rc = at3d::state_init_solution(st->s, msg);
  if (rc) set_err(err, errlen, msg);
  return rc;
}

int at3d_init_solution_from(at3d_state* st, at3d_state* src, char* err, size_t errlen) {
  if (!st || !src) {
    set_err(err, errlen, "null argument");
    return AT3D_ERR_ARG;
  }
  std::string msg;
  const int rc = at3d::state_init_solution_from(st->s, src->s, msg);
  if (rc) set_err(err, errlen, msg);
  return rc;
}

int at3d_solve(at3d_state* st, int32_t maxiter, int32_t verbose, float* solcrit_out,
               int32_t* iters_out, char* err, size_t errlen) {
  if (!st) {
    set_err(err, errlen, "null state");
    return AT3D_ERR_ARG;
  }
  std::string msg;
  int rc = at3d::state_solve(st->s, maxiter, verbose, msg);
  if (solcrit_out) *solcrit_out = st->s.solcrit;
  if (iters_out) *iters_out = st->s.iters;
  if (rc) set_err(err, errlen, msg);
  return rc;
}

int at3d_state_get(at3d_state* st, const char* name, void* dst, size_t nbytes, char* err,
                   size_t errlen) {
  if (!st || !name || !dst) {
    set_err(err, errlen, "null argument");
    return AT3D_ERR_ARG;
  }
  at3d::State& s = st->s;
  const std::string k(name);
  const size_t npts = s.npts;
  auto get1d = [&](const void* src, size_t bytes) -> int {
    if (nbytes < bytes) {
      set_err(err, errlen, "destination buffer too small");
      return AT3D_ERR_ARG;
    }
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s.stream);
    if (e != hipSuccess) {
      set_err(err, errlen, hipGetErrorString(e));
      return AT3D_ERR_DEVICE;
    }
    return AT3D_OK;
  };
  if (hipStreamSynchronize(s.stream) != hipSuccess) {
    set_err(err, errlen, "stream synchronize failed");
    return AT3D_ERR_DEVICE;
  }
  if (k == "source" || k == "radiance" || k == "delsource") {
    // [nstokes][npts][nlm]: the I plane, then Q and U of a polarized state
    const float* src = k == "source" ? s.S : (k == "radiance" ? s.R : s.DS);
    const size_t one = npts * (size_t)s.nlm * sizeof(float);
    if (nbytes < one * s.nst) {
      set_err(err, errlen, "destination buffer too small");
      return AT3D_ERR_ARG;
    }
    for (int q = 0; q < s.nst; q++) {
      hipError_t e = hipMemcpy2DAsync(static_cast<char*>(dst) + q * one, s.nlm * sizeof(float), src + q * s.psh,
                                      s.ld * sizeof(float), s.nlm * sizeof(float), npts, hipMemcpyDeviceToHost,
                                      s.stream);
      if (e != hipSuccess) {
        set_err(err, errlen, hipGetErrorString(e));
        return AT3D_ERR_DEVICE;
      }
    }
    if (hipStreamSynchronize(s.stream) != hipSuccess) {
      set_err(err, errlen, "stream synchronize failed");
      return AT3D_ERR_DEVICE;
    }
    return AT3D_OK;
  }
  if (k == "do_radiance" || k == "do_source") {
    // device layout [column][padded point] -> [reference ordinate][point]
    const size_t nang = s.nang;
    if (nbytes < nang * npts * sizeof(float)) {
      set_err(err, errlen, "destination buffer too small");
      return AT3D_ERR_ARG;
    }
    std::vector<float> tmp((size_t)s.nptsp * s.ldo);
    hipError_t e = hipMemcpyAsync(tmp.data(), k == "do_radiance" ? s.dorad : s.dosrc,
                                  tmp.size() * sizeof(float), hipMemcpyDeviceToHost, s.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s.stream);
    if (e != hipSuccess) {
      set_err(err, errlen, hipGetErrorString(e));
      return AT3D_ERR_DEVICE;
    }
    float* out = static_cast<float*>(dst);
    const bool level_order = k == "do_source" && s.dosrc_level_order;
    for (size_t k2 = 0; k2 < nang; k2++) {
      const size_t c = s.col_of_ref[k2];
      if (level_order) {
        int o = 0;
        while (o < 7 && (int)c >= s.oct_col_base[o] + s.oct_nord[o]) o++;
        const int32_t* pos = s.oct_pos_h.data() + (size_t)o * npts;
        for (size_t pnt = 0; pnt < npts; pnt++) out[k2 * npts + pnt] = tmp[c * s.nptsp + pos[pnt]];
      } else {
        for (size_t pnt = 0; pnt < npts; pnt++) out[k2 * npts + pnt] = tmp[c * s.nptsp + pnt];
      }
    }
    return AT3D_OK;
  }
  // the current grid (host copies in the state's plan)
  const at3d::Plan& p = *s.plan;
  if (k == "total_ext" || k == "extinct" || k == "albedo") {
    std::string msg;
    const int rc = at3d::ensure_host_medium(p, msg);
    if (rc) {
      set_err(err, errlen, msg);
      return rc;
    }
  }
  if (k == "gridpos") return copy_out(p.gridpos, dst, nbytes, err, errlen);
  if (k == "gridptr") return copy_out(p.gridptr, dst, nbytes, err, errlen);
  if (k == "neighptr") return copy_out(p.neighptr, dst, nbytes, err, errlen);
  if (k == "cellflags") return copy_out(p.cellflags, dst, nbytes, err, errlen);
  if (k == "total_ext") return copy_out(p.total_ext, dst, nbytes, err, errlen);
  if (k == "extinct") return copy_out(p.extinct, dst, nbytes, err, errlen);
  if (k == "albedo") return copy_out(p.albedo, dst, nbytes, err, errlen);
  if (k == "treeptr") {
    if (!p.treeptr.empty()) return copy_out(p.treeptr, dst, nbytes, err, errlen);
    return copy_out(std::vector<int32_t>((size_t)p.ncells * 2, -1), dst, nbytes, err, errlen);
  }
  // the state's own per-point medium (uploaded from the plan or made on the device)
  const size_t ml1 = p.ml + 1;
  if (k == "point_ext") return get1d(s.ext, npts * sizeof(float));
  if (k == "point_alb") return get1d(s.alb, npts * sizeof(float));
  if (k == "point_chi") return get1d(s.chi, npts * ml1 * sizeof(float));
  if (k == "point_tc") return get1d(s.tc, npts * ml1 * sizeof(float));
  if (k == "point_tms") return get1d(s.tms, npts * ml1 * sizeof(float));   // times the direct flux
  if (k == "point_radiance0") return get1d(s.R0, npts * 4 * sizeof(float));
  if (k == "point_ss_off") return get1d(s.ss_off, (npts + 1) * sizeof(int32_t));
  if (k == "fluxes") return get1d(s.fluxes, 2 * npts * sizeof(float));
  if (k == "dirflux") return get1d(s.dirflux, npts * sizeof(float));
  if (k == "ns") return get1d(s.ns, npts * sizeof(int));
  if (k == "nr") return get1d(s.nr, npts * sizeof(int));
  if (k == "point_ss_iph" || k == "point_ss_coef") {
    int32_t n = 0;
    if (hipMemcpy(&n, s.ss_off + npts, sizeof n, hipMemcpyDeviceToHost) != hipSuccess) {
      set_err(err, errlen, "device copy failed");
      return AT3D_ERR_DEVICE;
    }
    return k == "point_ss_iph" ? get1d(s.ss_iph, (size_t)n * 4) : get1d(s.ss_coef, (size_t)n * 4);
  }
  set_err(err, errlen, "unknown state array '" + k + "'");
  return AT3D_ERR_ARG;
}

int at3d_render(at3d_state* st, int64_t nrays, const float* x, const float* y, const float* z,
                const double* mu, const double* phi, int32_t single_scatter, int32_t nosurface,
                float* out, char* err, size_t errlen) {
  if (!st || (nrays > 0 && (!x || !y || !z || !mu || !phi || !out))) {
    set_err(err, errlen, "null argument");
    return AT3D_ERR_ARG;
  }
  std::string msg;
  int rc;
  try {
    rc = at3d::state_render(st->s, nrays, x, y, z, mu, phi, single_scatter, nosurface, out, msg);
  } catch (const std::bad_alloc&) {
    rc = AT3D_ERR_MEMORY;
    msg = "out of host memory in render";
  }
  if (rc) set_err(err, errlen, msg);
  return rc;
}

int at3d_state_load(at3d_state* st, const int32_t* shptr, const float* source,
                    const int32_t* rshptr, const float* radiance, const float* fluxes, char* err,
                    size_t errlen) {
  if (!st || !shptr || !rshptr || !radiance || !fluxes || !source) {
    set_err(err, errlen, "null argument");
    return AT3D_ERR_ARG;
  }
  std::string msg;
  int rc;
  try {
    rc = at3d::state_load(st->s, shptr, source, rshptr, radiance, fluxes, msg);
  } catch (const std::bad_alloc&) {
    rc = AT3D_ERR_MEMORY;
    msg = "out of host memory loading a solution";
  }
  if (rc) set_err(err, errlen, msg);
  return rc;
}

int at3d_state_set_grid(at3d_state* st, int32_t npts, int32_t ncells, const float* gridpos,
                        const int32_t* gridptr, const int32_t* neighptr, const int32_t* treeptr,
                        const int8_t* cellflags, char* err, size_t errlen) {
  if (!st || !gridpos || !gridptr || !neighptr || !treeptr || !cellflags || npts <= 0 || ncells <= 0) {
    set_err(err, errlen, "null argument");
    return AT3D_ERR_ARG;
  }
  std::string msg;
  int rc;
  try {
    rc = at3d::state_set_grid(st->s, npts, ncells, gridpos, gridptr, neighptr, treeptr, cellflags, msg);
  } catch (const std::bad_alloc&) {
    rc = AT3D_ERR_MEMORY;
    msg = "out of host memory setting a grid";
  }
  if (rc) set_err(err, errlen, msg);
  return rc;
}

int at3d_state_save(at3d_state* st, int32_t* shptr, int32_t* rshptr, float* source,
                    int64_t source_cap, float* radiance, int64_t radiance_cap, float* fluxes,
                    char* err, size_t errlen) {
  if (!st || !shptr || !rshptr) {
    set_err(err, errlen, "null argument");
    return AT3D_ERR_ARG;
  }
  std::string msg;
  int rc;
  try {
    rc = at3d::state_save(st->s, shptr, rshptr, source, source_cap, radiance, radiance_cap, fluxes,
                          msg);
  } catch (const std::bad_alloc&) {
    rc = AT3D_ERR_MEMORY;
    msg = "out of host memory saving a solution";
  }
  if (rc) set_err(err, errlen, msg);
  return rc;
}

int at3d_state_sh_terms(at3d_state* st, int32_t nshout, float* shterms, float* netfluxdiv,
                        char* err, size_t errlen) {
  if (!st) {
    set_err(err, errlen, "null state");
    return AT3D_ERR_ARG;
  }
  std::string msg;
  const int rc = at3d::state_sh_terms(st->s, nshout, shterms, netfluxdiv, msg);
  if (rc) set_err(err, errlen, msg);
  return rc;
}

int at3d_render_noalloc(at3d_state* st, int64_t nrays, const float* x, const float* y,
                        const float* z, const double* mu, const double* phi, int32_t single_scatter,
                        int32_t nosurface, float* out, char* err, size_t errlen) {
  if (!st || (nrays > 0 && (!x || !y || !z || !mu || !phi || !out))) {
    set_err(err, errlen, "null argument");
    return AT3D_ERR_ARG;
  }
  std::string msg;
  int rc;
  try {
    rc = at3d::state_render_noalloc(st->s, nrays, x, y, z, mu, phi, single_scatter, nosurface, out,
                                    msg);
  } catch (const std::bad_alloc&) {
    rc = AT3D_ERR_MEMORY;
    msg = "out of host memory in render";
  }
  if (rc) set_err(err, errlen, msg);
  return rc;
}

int64_t at3d_render_scratch_bytes(void) { return (int64_t)at3d::RENDER_SCRATCH_BYTES; }

int at3d_levis_gradient_jacobian(at3d_state* st, const at3d_deriv_desc* deriv, int64_t nrays,
                                 const float* x, const float* y, const float* z, const double* mu,
                                 const double* phi, int64_t npix, const int32_t* rays_per_pixel,
                                 const double* ray_weights, const double* stokes_weights,
                                 const float* measurements, const double* uncertainties,
                                 const char* costfunc, int32_t exact_single_scatter, int32_t njac,
                                 const int32_t* jacobian_ptr, float* jacobian, double* grad,
                                 double* cost, float* pixels, char* err, size_t errlen) {
  if (!st || !deriv || !x || !y || !z || !mu || !phi || !rays_per_pixel || !ray_weights ||
      !stokes_weights || !measurements || !uncertainties || !grad || !cost || !pixels ||
      !deriv->partder || !deriv->dext || !deriv->dalb || (njac > 0 && (!jacobian_ptr || !jacobian))) {
    set_err(err, errlen, "null argument");
    return AT3D_ERR_ARG;
  }
  std::string msg;
  int rc;
  try {
    rc = at3d::state_levis_gradient(st->s, *deriv, nrays, x, y, z, mu, phi, npix, rays_per_pixel,
                                    ray_weights, stokes_weights, measurements, uncertainties,
                                    costfunc, exact_single_scatter, njac, jacobian_ptr, jacobian,
                                    grad, cost, pixels, msg);
  } catch (const std::bad_alloc&) {
    rc = AT3D_ERR_MEMORY;
    msg = "out of host memory in the gradient";
  }
  if (rc) set_err(err, errlen, msg);
  return rc;
}

int at3d_levis_gradient(at3d_state* st, const at3d_deriv_desc* deriv, int64_t nrays,
                        const float* x, const float* y, const float* z, const double* mu,
                        const double* phi, int64_t npix, const int32_t* rays_per_pixel,
                        const double* ray_weights, const double* stokes_weights,
                        const float* measurements, const double* uncertainties,
                        const char* costfunc, int32_t exact_single_scatter, double* grad,
                        double* cost, float* pixels, char* err, size_t errlen) {
  return at3d_levis_gradient_jacobian(st, deriv, nrays, x, y, z, mu, phi, npix, rays_per_pixel,
                                      ray_weights, stokes_weights, measurements, uncertainties,
                                      costfunc, exact_single_scatter, 0, nullptr, nullptr, grad,
                                      cost, pixels, err, errlen);
}

int at3d_state_set_option(at3d_state* st, const char* key, int32_t value, char* err,
                          size_t errlen) {
  if (!st || !key) {
    set_err(err, errlen, "null argument");
    return AT3D_ERR_ARG;
  }
  const std::string k(key);
  if (k == "sweep_general") {
    st->s.force_general = value ? 1 : 0;
    return AT3D_OK;
  }
  if (k == "sweep_graph") {
    st->s.use_graph = value ? 1 : 0;
    return AT3D_OK;
  }
  if (k == "sweep_ring") {
    st->s.use_ring = value < 0 ? 0 : (value > 2 ? 2 : value);
    return AT3D_OK;
  }
  if (k == "render_correct_interpolate") {
    st->s.render_correct_interp = value ? 1 : 0;
    return AT3D_OK;
  }
  if (k == "sweep_ring4") {
    std::string e;
    const int rc = state_set_ring4(st->s, value != 0, e);
    if (rc) set_err(err, errlen, e.c_str());
    return rc;
  }
  if (k == "iteration_graph") {
    st->s.use_iter_graph = value ? 1 : 0;
    return AT3D_OK;
  }
  if (k == "max_subgrid_intervals") {
    st->s.max_subgrid = value > 0 ? value : 0;
    return AT3D_OK;
  }
  set_err(err, errlen, "unknown option '" + k + "'");
  return AT3D_ERR_ARG;
}

double at3d_state_timing(const at3d_state* st, const char* key) {
  if (!st || !key) return -1.0;
  const at3d::State& s = st->s;
  const std::string k(key);
  if (k == "truncation") return s.tsum[0];
  if (k == "sh_to_do") return s.tsum[1];
  if (k == "sweep") return s.tsum[2];
  if (k == "do_to_sh") return s.tsum[3];
  if (k == "source") return s.tsum[4];
  if (k == "sweep_levels")
    return (double)s.nlevels_max;
  if (k == "total") return s.tsum[0] + s.tsum[1] + s.tsum[2] + s.tsum[3] + s.tsum[4] + s.tsum[5];
  if (k == "sweep_launches") return s.nsweep_launch;
  if (k == "sweep_kernel") return s.tsum[6];
  if (k == "device_bytes") return (double)s.device_bytes;
  if (k == "sweep_ring")
    return (s.ring_ok && s.use_ring && !s.force_general && (s.use_ring == 2 || !s.ring_chain)) ? 1.0 : 0.0;
  if (k == "sweep_ring_lds") return (double)(s.ring_full ? s.ring_lds_full : s.ring_lds);
  if (k == "sweep_ring_slots") return s.slots;
  if (k == "sweep_ring4") return s.ring4;
  if (k == "sweep_ring_full") return s.ring_full;
  if (k == "sweep_sectors") return s.nsectors;
  if (k == "sweep_general_points") return s.ngeneral;
  if (k == "iters") return s.iters_last;
  if (k == "render") return s.render_ms;
  if (k == "render_rays") return s.render_ray_ms;
  if (k == "iteration_graph_active") return s.last_solve_graph ? 1.0 : 0.0;
  if (k == "adjoint_rays") return s.adjoint_ms;
  if (k == "gradient") return s.grad_ms;
  if (k == "splitcrit") return s.splitcrit;
  return -1.0;
}

int64_t at3d_state_size(const at3d_state* st, const char* key) {
  if (!st || !key || !st->s.plan) return -1;
  const at3d::Plan& p = *st->s.plan;
  const std::string k(key);
  if (k == "npts") return p.npts;
  if (k == "ncells") return p.ncells;
  if (k == "nbpts") return p.nbpts;
  if (k == "nbcells") return p.nbcells;
  if (k == "ntop") return (int64_t)p.bctop.size();
  if (k == "nbot") return (int64_t)p.bcbot.size();
  if (k == "refined") return p.treeptr.empty() ? 0 : 1;
  return -1;
}

}  // extern "C"
