// Device-resident state of one solve.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "devmem.h"
#include "host_plan.h"
#include "kernels.h"

namespace at3d {

// Device scratch bytes of the allocation-free render (at3d_render_noalloc):
// the direction source of every grid point for every distinct ray direction
// plus the rays; 64 MiB holds a 64 x 64 x 40 grid with up to 64 directions
// and 2^20 rays.
constexpr size_t RENDER_SCRATCH_BYTES = (size_t)64 << 20;

struct State {
  const Plan* plan = nullptr;
  // adaptive cell splitting: the refined grid's plan, owned by the state
  // (null while the state is on its base grid), and the plan it started from
  std::shared_ptr<Plan> own_plan;
  const Plan* base_plan = nullptr;
  int32_t* treeptr = nullptr;   // device [ncells][2] of a refined grid, else null
  float splitcrit = 0.0f;       // SPLITCRIT of the last splitting test
  float albmax = 0.0f;          // the largest scattering albedo (a single iteration below SOLACC)
  float* beam_ext = nullptr;    // device medium: the direct beam's property-grid extinction
  std::vector<int32_t> medium_cls_h;   // device medium: source classes and total extinction
  std::vector<float> medium_ext_h;     // of every point (for the ring layout)
  int device = 0;   // HIP device the state was created on
  // polarization (NSTOKES = 3): S, R, DS hold the I, Q, U planes psh floats
  // apart, dosrc / dorad theirs pdo floats apart; the Q / U transform tables
  // ([2 ldo][2 ld], by device column), the a2 / a3 / b1 source terms and the
  // b1 single-scatter terms; the SH->DO tiles of the Q and U rows
  int nst = 1;
  size_t psh = 0, pdo = 0;
  float *Yqu = nullptr, *Wqu = nullptr, *chip = nullptr, *tms5 = nullptr;
  int4* sh_tiles_qu = nullptr;
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
  int lo_base_h[8] = {};   // per-octant base into lvl_off
  int nlevels_max = 0;
  // transforms
  float *Y = nullptr, *W = nullptr, *ysun = nullptr;
  int* lofj = nullptr;
  // optics
  float *ext = nullptr, *alb = nullptr, *chi = nullptr, *tc = nullptr, *dirflux = nullptr, *f0 = nullptr;
  // SH fields [nptsp][ld]
  float *S = nullptr, *R = nullptr, *DS = nullptr;
  float* R0 = nullptr;   // [npts][4] two-stream initial radiance terms
  double* host_dots = nullptr;   // pinned: four dot products, then the error flag
  int *ns = nullptr, *nr = nullptr, *nos = nullptr;
  int* nr_next = nullptr;   // NR for the next iteration, written by the source pass
  bool nr_next_fixsh = false;   // nr_next was made with FIXSH on
  // DO fields [ldo][nptsp] (ordinate-major)
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
  int32_t* lvl_off = nullptr;
  int8_t* col_oct = nullptr;
  // level-ordered sweep (k_sweep_ring)
  bool ring_ok = false;        // ring fits in LDS and records are built
  int use_ring = 1;            // option: sweep with k_sweep_ring when ring_ok (1: unless ring_chain, 2: always)
  bool ring_chain = false;     // the levels are chains (under 4 entries per level in some octant)
  int slots = 0, sector_points = 8, gld = 0, ngeneral = 0, nsectors = 0;
  int ring_nlev[8] = {}, sec_base[8] = {};
  size_t ring_lds = 0;         // dynamic LDS bytes of k_sweep_ring's fast instantiation (0: unused)
  size_t ring_lds_full = 0;    // and of its full instantiation
  int32_t *oct_perm = nullptr, *oct_pos = nullptr, *top_index = nullptr;
  int32_t *ring_code = nullptr, *ring_cell = nullptr, *ring_level = nullptr, *ring_off = nullptr;
  int8_t* ring_staged = nullptr;
  int32_t *sec_off = nullptr, *sec_p0 = nullptr;
  uint4* sec_slot = nullptr;
  std::vector<int32_t> oct_pos_h;   // [8][npts]
  uint4* rec = nullptr;
  float4* geo = nullptr;       // [nang][ncls] step classes
  int2* geo_rep = nullptr;     // [8][ncls] class representatives
  float* ext_pos = nullptr;    // [8][nptsp] extinction in ring order
  int ncls = 0;
  unsigned long long nz_magic = 0;
  float cls_tol = 2e-5f;
  uint8_t* gflag = nullptr;
  uint8_t* face_ok = nullptr;   // [8][npts] one-cell faces of each point (k_face_mask)
  int* col_general = nullptr;
  std::vector<int32_t> col_general_h;   // columns with multi-cell traces
  int ring_full = 0;                    // every column needs the full ring instantiation
  int ring4 = 0;                        // the ring holds radiances alone (4-byte slots)
  int force_ring4 = 0;                  // option: 4-byte slots even where pairs fit
  int* ngeneral_d = nullptr;
  int4* sh_tiles = nullptr;
  int n_sh_tiles = 0;
  bool dosrc_level_order = false;   // layout of dosrc written by the last SH->DO
  // options
  int force_general = 0;  // 1: every point through the general characteristic trace
  int use_graph = 1;      // replay the sweep level launches from a HIP graph
  hipGraph_t sweep_graph = nullptr;
  hipGraphExec_t sweep_exec = nullptr;
  int graph_general = -1, graph_ring = -1;
  // iterations replayed from captured graphs under device control (IterCtl):
  // ITER_BATCH iterations per graph, one graph per NR buffer parity
  static const int ITER_BATCH = 4;
  IterCtl* ctl = nullptr;      // device
  IterCtl* ctl_h = nullptr;    // pinned host mirror
  int use_iter_graph = 1;      // option "iteration_graph"
  int render_correct_interp = 1;   // option "render_correct_interpolate"
  int* nr_buf0 = nullptr;      // the first NR buffer (the graphs' parity reference)
  hipGraph_t iter_graph[2] = {};
  hipGraphExec_t iter_exec[2] = {};
  int iter_key[2] = {-1, -1};  // sweep configuration each graph was captured with
  bool last_solve_graph = false;   // the last solve call ran its iterations from graphs
  // iteration state
  bool initialized = false;
  int iters = 0, iters_last = 0, nsweep_launch = 0;
  float solcrit = 1.0f, deljdot = 0, deljold = 0, deljnew = 0, jnorm = 0;
  static const int NEV = 8;
  hipEvent_t ev[NEV] = {};
  hipEvent_t ev_sweep[4] = {};  // around the two ring sweep launches of an iteration
  hipEvent_t ev_ray[2] = {};    // around the ray kernel of the last render / gradient call
  double tsum[8] = {};  // ms per kernel class over the last solve
  std::vector<std::pair<void*, size_t>> allocs;   // device blocks and their sizes
  size_t device_bytes = 0;   // bytes of the allocations in `allocs`
  // rendering
  float *tms = nullptr, *ss_coef = nullptr, *phasetab = nullptr, *bcrad_render = nullptr;
  int32_t *ss_off = nullptr, *ss_iph = nullptr, *bot_index = nullptr;
  // per-call render buffers (grown on demand, freed with the state)
  size_t rbuf_bytes = 0;
  void* rbuf = nullptr;
  size_t gbuf_bytes = 0;   // per-call gradient buffers
  void* gbuf = nullptr;
  int max_subgrid = 0;     // extra sub-step intervals a gradient ray may record
  double grad_ms = 0.0;
  double render_ms = 0.0;
  double render_ray_ms = 0.0;   // k_render alone in the last render call
  double adjoint_ms = 0.0;      // k_adjoint_rays alone in the last gradient call
};

struct DirTables {
  int ndir = 0, ldd = 0;
  std::vector<int32_t> dir_of_ray;                  // [nrays]
  std::vector<std::pair<double, double>> dirs;      // (mu, phi) of each direction
  std::vector<float> Yd, G, ssd;                    // [ldd][ld], [ndir][ml+1], [ndir][numphase]
  // polarized: the Q / U rows [2 ldd][2 ld] (Q of direction d: elements 2 and
  // 5; U at ldd + d: elements 6 and 3), the Q and U truncation factors
  // [2][ndir][ml+1] after G, and the rotated I-Q phase values [ndir][numphase]
  std::vector<float> Yqu, ssq, ssu;
  std::vector<int32_t> sa_j;                        // scattering-angle table index (1-based lower)
  std::vector<float> sa_f;                          // and interpolation fraction
};
int direction_tables(const State& s, int64_t nrays, const double* mu, const double* phi, DirTables& t,
                     std::string& err);

int state_levis_gradient(State& s, const at3d_deriv_desc& dd, int64_t nrays, const float* x,
                         const float* y, const float* z, const double* mu, const double* phi,
                         int64_t npix, const int32_t* rays_per_pixel, const double* ray_weights,
                         const double* stokes_weights, const float* measurements,
                         const double* uncertainties, const char* costfunc, int ess, int32_t njac,
                         const int32_t* jacptr, float* jacobian, double* grad, double* cost,
                         float* pixels, std::string& err);

int state_render(State& s, int64_t nrays, const float* x, const float* y, const float* z,
                 const double* mu, const double* phi, int single_scatter, int nosurface,
                 float* out, std::string& err);


int state_render_noalloc(State& s, int64_t nrays, const float* x, const float* y, const float* z,
                         const double* mu, const double* phi, int single_scatter, int nosurface,
                         float* out, std::string& err);

int state_load(State& s, const int32_t* shptr, const float* source, const int32_t* rshptr,
               const float* radiance, const float* fluxes, std::string& err);
int state_save(State& s, int32_t* shptr, int32_t* rshptr, float* source, int64_t source_cap,
               float* radiance, int64_t radiance_cap, float* fluxes, std::string& err);
int state_sh_terms(State& s, int nshout, float* shterms, float* netfluxdiv, std::string& err);

int state_create(const Plan& p, State& s, std::string& err);
// Moves a state on a refined grid back to its base grid (no solution: the
// next call initialises or loads one).
int state_reset_grid(State& s, std::string& err);
// Moves a state to the grid of a saved solution (the base grid or a refined
// one); no solution until the next load or initialisation.
int state_set_grid(State& s, int npts, int ncells, const float* gridpos, const int32_t* gridptr,
                   const int32_t* neighptr, const int32_t* treeptr, const int8_t* cellflags, std::string& err);
size_t plan_footprint_bytes(const Plan& p);
void state_free(State& s);
int state_init_solution(State& s, std::string& err);
int state_init_solution_from(State& s, State& src, std::string& err);
int state_solve(State& s, int maxiter, int verbose, std::string& err);
// Rebuilds the level-ordered sweep's layout with a ring of radiances alone
// (on) or of (radiance, source) pairs where they fit (off); the captured
// graphs are dropped.  The solution is kept.
int state_set_ring4(State& s, bool on, std::string& err);

}  // namespace at3d
