// Host-side launch wrappers of the gfx950 kernels (one translation unit each).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "shdom_geom.h"

namespace at3d {

// ---- transforms.hip: FP32 MFMA SH <-> DO transforms ----------------------
// dosrc[p][c] = max(0, sum_J S[p][J] * Yc[c][J]) * ext[p]
// S [mp][ld], Yc [ncp][ld], dosrc [mp][ldo]; mp, ncp multiples of 64; ld multiple of 16.
void launch_sh_to_do(hipStream_t st, const float* S, const float* Yc, const float* ext,
                     float* dosrc, int mp, int ncp, int ld, int ldo, int npts, int nang);
// R[p][J] = (J < nr[p]) ? sum_c dorad[p][c] * Wc[c][J] : 0
// dorad [mp][ldo], Wc [ldo][ld] (rows beyond nang zero), R [mp][ld].
void launch_do_to_sh(hipStream_t st, const float* dorad, const float* Wc, const int* nr, float* R,
                     int mp, int ld, int ldo, int npts);

// ---- sweep.hip: discrete-ordinate sweeps --------------------------------
struct SweepArgs {
  GridTopo g;
  const OrdGeom* ordg;       // [ncols] device column order
  const float* skyrad;       // [ncols]
  const float* bcrad_bot;    // [nbot]
  const int32_t* bctop;      // [ntop]
  const int32_t* bcbot;      // [nbot]
  int ntop, nbot;
  const int8_t* preset;      // [npts] bit0 top point, bit1 bottom point
  const int32_t* rank;       // [8][npts]
  const int32_t* level;      // [8][npts]
  const int32_t* lvl_pts;    // concatenated over octants (~pt = general trace)
  const int32_t* lvl_cell;
  const float* ext;          // [npts]
  const float* srcext;       // [npts][ldo]
  float* rad;                // [npts][ldo]
  const float* xgrid;
  const float* ygrid;
  const float* zgrid;
  int nx1, ny1, nz;
  int ldo, npts;
  int force_general;
  int* errflag;
};

// Items of one level launch: up to four octants of one hemisphere, each
// contributing (points of this level) x (ordinates of the octant).
struct LevelDesc {
  int item_off[5];
  int pts_off[4];   // into the concatenated level point list
  int nord[4];
  int col_base[4];
  int oct[4];
  int nslot;
  int L;
};

void launch_sweep_preset(hipStream_t st, const SweepArgs& a, bool up, int col0, int ncol);
void launch_sweep_level(hipStream_t st, const SweepArgs& a, const LevelDesc& d);
// Lambertian surface radiance from the downwelling flux (reference ordinate order).
void launch_bottom_bc(hipStream_t st, const float* rad, int ldo, const int32_t* bcbot, int nbot,
                      const int32_t* col_of_ref, const float* fluxw, int ndown,
                      const float* dirflux, float albpi, float* bcrad);
void launch_fluxes(hipStream_t st, const float* rad, int ldo, const int32_t* col_of_ref,
                   const float* fluxw, int ndown, int nang, int npts, float* fluxes);

// ---- source.hip: truncation, source function, acceleration ---------------
struct SourceArgs {
  const float* R;
  const int* nr;
  const float* alb;
  const float* chi;   // [npts][ml+1]
  const float* f0;    // DIRFLUX * SECMU0
  const float* ysun;  // [nlm]
  const int* lofj;    // [nlm]
  float* S;
  float* DS;
  int* ns;
  int* nos;
  double* partial;    // [nblocks][4]
  int ld, npts, ml, mm, nlm;
  float shacc;
  int first, accel, fixsh;
};
constexpr int SRC_PTS_PER_BLOCK = 16;
void launch_radiance_truncation(hipStream_t st, const float* R, int ld, const float* tc,
                                const int* ns, const int* lofj, int* nr, int npts, int ml, int mm,
                                int nlm, float radmin, int fixsh, int highorder);
void launch_compute_source(hipStream_t st, const SourceArgs& a, int nblocks);
void launch_reduce_dots(hipStream_t st, const double* partial, int nblocks, double* out);
void launch_accelerate(hipStream_t st, float* S, const float* DS, const int* ns, const int* nos,
                       int ld, int npts, float a);

// ---- render.hip: rays ----------------------------------------------------
struct RayArgs {
  GridTopo g;
  const float* xgrid;
  const float* ygrid;
  const float* zgrid;
  int nx, ny, nz, nx1, ny1, nyc;
  int open_x, open_y;
  const float* ext;        // [npts]
  const float* dsrc;       // [npts][ldd]
  int ldd;
  const int32_t* bot_index;  // [npts] bottom boundary index or -1
  const float* bcrad_bot;    // [nbot]
  const float* camx;
  const float* camy;
  const float* camz;
  const double* cammu;
  const double* camphi;
  const int32_t* dir_of_ray;
  const float* topval;       // [nrays] downwelling radiance at the domain top
  int nrays;
  double tautol, transcut;
  int nosurface;
  float* out;
  int* errflag;
};
void launch_dir_source_finish(hipStream_t st, const float* raw, int ldd, int ndir, const float* ext,
                              const int* ns, const int* lofj, const float* tms, const float* G,
                              const int32_t* ss_off, const int32_t* ss_iph, const float* ss_coef,
                              const float* ssd, int numphase, int ml, int npts, int single_scatter,
                              int tms_on, float* dsrc);
void launch_render(hipStream_t st, const RayArgs& a);
void launch_render_bottom(hipStream_t st, const float* fluxes, const float* dirflux,
                          const int32_t* bcbot, int nbot, float albpi, float* bcrad);
// raw[p][d] = sum_J S[p][J] * Yd[d][J] (no clamp, no extinction)
void launch_sh_dot(hipStream_t st, const float* S, const float* Yd, float* raw, int mp, int ndp,
                   int ld, int ldd, int npts, int ndir);

}  // namespace at3d
