// Device-resident state of one solve.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "host_plan.h"
#include "kernels.h"

namespace at3d {

struct State {
  const Plan* plan = nullptr;
  hipStream_t stream = nullptr;
  int npts = 0, nlm = 0, nang = 0;
  int ld = 0;      // row stride of SH arrays (nlm rounded up to 64)
  int nptsp = 0;   // points rounded up to 64
  int ldo = 0;     // row stride of DO arrays (nang rounded up to 64)
  int ndown = 0;   // number of downwelling ordinates
  int nsrc_blocks = 0;
  // ordinate columns: grouped by octant, reference order within an octant
  std::vector<int32_t> col_of_ref, ref_of_col;
  int oct_col_base[8] = {}, oct_nord[8] = {}, oct_pts_base_h[8] = {};
  std::vector<LevelDesc> levels_down, levels_up;
  // transforms
  float *Y = nullptr, *W = nullptr, *ysun = nullptr;
  int* lofj = nullptr;
  // optics
  float *ext = nullptr, *alb = nullptr, *chi = nullptr, *tc = nullptr, *dirflux = nullptr, *f0 = nullptr;
  // SH fields [nptsp][ld]
  float *S = nullptr, *R = nullptr, *DS = nullptr, *R0 = nullptr;
  int *ns = nullptr, *nr = nullptr, *nos = nullptr;
  // DO fields [nptsp][ldo]
  float *dosrc = nullptr, *dorad = nullptr;
  float* fluxes = nullptr;
  double *partial = nullptr, *dots = nullptr;
  int* errflag = nullptr;
  // sweep data
  float *xgrid = nullptr, *ygrid = nullptr, *zgrid = nullptr;
  float* gridpos = nullptr;
  int32_t *gridptr = nullptr, *neighptr = nullptr;
  int8_t* cellflags = nullptr;
  OrdGeom* ordg = nullptr;
  float *skyrad = nullptr, *fluxw = nullptr, *bcrad_bot = nullptr;
  int32_t *col_of_ref_d = nullptr;
  int32_t *bctop = nullptr, *bcbot = nullptr;
  int8_t* preset = nullptr;
  int32_t *rank = nullptr, *level = nullptr, *lvl_pts = nullptr, *lvl_cell = nullptr;
  // options
  int force_general = 0;  // 1: every point through the general characteristic trace
  int use_graph = 1;      // replay the sweep level launches from a HIP graph
  hipGraph_t sweep_graph = nullptr;
  hipGraphExec_t sweep_exec = nullptr;
  int graph_general = -1;
  // iteration state
  bool initialized = false;
  int iters = 0, iters_last = 0, nsweep_launch = 0;
  float solcrit = 1.0f, deljdot = 0, deljold = 0, deljnew = 0, jnorm = 0;
  static const int NEV = 8;
  hipEvent_t ev[NEV] = {};
  double tsum[8] = {};  // ms per kernel class over the last solve
  std::vector<void*> allocs;
  // rendering
  float *tms = nullptr, *ss_coef = nullptr, *phasetab = nullptr, *bcrad_render = nullptr;
  int32_t *ss_off = nullptr, *ss_iph = nullptr, *bot_index = nullptr;
  // per-call render buffers (grown on demand, freed with the state)
  size_t rbuf_bytes = 0;
  void* rbuf = nullptr;
  double render_ms = 0.0;
};

int state_render(State& s, int64_t nrays, const float* x, const float* y, const float* z,
                 const double* mu, const double* phi, int single_scatter, int nosurface,
                 float* out, std::string& err);

int state_create(const Plan& p, State& s, std::string& err);
void state_free(State& s);
int state_init_solution(State& s, std::string& err);
int state_solve(State& s, int maxiter, int verbose, std::string& err);

}  // namespace at3d
