// Device-resident state of one solve.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "host_plan.h"

namespace at3d {

struct State {
  const Plan* plan = nullptr;
  hipStream_t stream = nullptr;
  int npts = 0, nlm = 0, nang = 0;
  int ld = 0;      // row stride of SH arrays (nlm rounded up to 64)
  int nptsp = 0;   // points rounded up to 64 (row stride of DO arrays)
  int nangp = 0;   // ordinates rounded up to 64
  int ndown = 0;   // number of downwelling ordinates (they come first)
  int nsrc_blocks = 0;
  // transforms
  float *Y = nullptr, *W = nullptr, *ysun = nullptr;
  int* lofj = nullptr;
  // optics
  float *ext = nullptr, *alb = nullptr, *chi = nullptr, *tc = nullptr, *dirflux = nullptr, *f0 = nullptr;
  // SH fields [nptsp][ld]
  float *S = nullptr, *R = nullptr, *DS = nullptr, *R0 = nullptr;
  int *ns = nullptr, *nr = nullptr, *nos = nullptr;
  // DO fields [nangp][nptsp]
  float *dosrc = nullptr, *dorad = nullptr;
  float* fluxes = nullptr;
  double *partial = nullptr, *dots = nullptr;
  int* errflag = nullptr;
  // sweep data
  float* gridpos = nullptr;
  int32_t *gridptr = nullptr, *neighptr = nullptr;
  int8_t* cellflags = nullptr;
  OrdGeom* ordg = nullptr;
  int32_t* ordoct = nullptr;
  float *skyrad = nullptr, *fluxw = nullptr, *bcrad_bot = nullptr;
  int32_t *bctop = nullptr, *bcbot = nullptr;
  int8_t* preset = nullptr;
  int32_t *rank = nullptr, *level = nullptr, *lvl_pts = nullptr, *lvl_cell = nullptr, *lvl_off = nullptr;
  int32_t *oct_lvl_base = nullptr, *oct_pts_base = nullptr, *oct_nlev = nullptr;
  // iteration state
  bool initialized = false;
  int iters = 0, iters_last = 0;
  float solcrit = 1.0f, deljdot = 0, deljold = 0, deljnew = 0, jnorm = 0;
  static const int NEV = 8;
  hipEvent_t ev[NEV] = {};
  double tsum[8] = {};  // ms per kernel class over the last solve
  std::vector<void*> allocs;
};

int state_create(const Plan& p, State& s, std::string& err);
void state_free(State& s);
int state_init_solution(State& s, std::string& err);
int state_solve(State& s, int maxiter, int verbose, std::string& err);

}  // namespace at3d
