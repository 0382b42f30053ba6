// Host-side preparation of a Levis-approximation gradient (CPU only).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "host_plan.h"

namespace at3d {

struct GradPlan {
  int numder = 0, maxpg = 0, dnumphase = 1;
  std::vector<int32_t> partder;                  // [numder] 0-based species
  std::vector<int32_t> doexact;                  // [numder]
  std::vector<float> dext, dalb, dphasewtp;      // [numder][maxpg]
  std::vector<int32_t> diphasep;                 // [numder][maxpg] 0-based row of dleg
  std::vector<float> dleg;                       // [dnumphase][nleg+1], divided by (2l+1)
  std::vector<float> dphasetab;                  // [dnumphase][nscatangle]
  std::vector<float> optinterpwt;                // [npts][8]
  std::vector<int32_t> interpptr;                // [npts][8] 0-based property point
  std::vector<float> dextm;                      // [numder][maxpg]
  std::vector<float> dalbm, dfj;                 // [numder][npts][8]
  // sun paths transposed to the property grid: for property point ib, the
  // (RTE point, d tau / d ext) pairs bp_ip/bp_w[bp_off[ib] .. bp_off[ib+1])
  std::vector<int32_t> bp_off, bp_ip;
  std::vector<float> bp_w;
};

// Property-grid -> RTE-grid interpolation, delta-M scaled derivatives
// (PREPARE_DERIV_INTERPS, COMPUTE_INTERP_WEIGHTS shdomsub4.f:2917-3169), the
// derivative phase table (PRECOMPUTE_PHASE_CHECK_GRAD :2493-2587) and the sun
// paths (MAKE_DIRECT_DERIVATIVE shdomsub5.f:1553-2003).
int build_grad_plan(const Plan& p, const at3d_deriv_desc& d, GradPlan& g, std::string& err);

}  // namespace at3d
