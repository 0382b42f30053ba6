// Host-side setup of one SHDOM solve (CPU only).  See host_plan.cpp.
#pragma once
#include <stdint.h>

#include <string>
#include <algorithm>
#include <array>
#include <memory>
#include <thread>
#include <vector>

#include "direct_beam.h"

#include "../../include/at3d_hip.h"
#include "medium_point.h"
#include "shdom_geom.h"

namespace at3d {

// Runs f(0..n-1) on up to 16 host threads (independent items only).
template <typename F>
inline void parallel_for(int n, F&& f) {
  const int nt = std::max(1, std::min(n, std::min(16, (int)std::thread::hardware_concurrency())));
  if (nt <= 1) {
    for (int i = 0; i < n; i++) f(i);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt);
  for (int t = 0; t < nt; t++)
    th.emplace_back([&, t]() {
      for (int i = t; i < n; i += nt) f(i);
    });
  for (std::thread& x : th) x.join();
}

struct Ordinate {
  int imu, iphi;      // 0-based zenith / azimuth index
  float mu, phi;
  float wtdo;
  OrdGeom g;          // trace-back geometry
  bool down;
};

struct OctantPlan {
  std::vector<int32_t> sweepord;   // [npts] (cell<<3)|corner in sweep order
  std::vector<int32_t> rank;       // [npts] position of each point in sweepord
  std::vector<int32_t> level;      // [npts] dependency level (0 = boundary preset)
  std::vector<int8_t> simple_pt;   // [npts] one-cell characteristic in a regular cell
  std::vector<int32_t> lvl_pts;    // non-preset points sorted by (level, rank);
                                   // ~point for points that need the general trace
  std::vector<int32_t> lvl_cell;   // their start cell
  std::vector<int32_t> lvl_off;    // [nlevels+1] offsets into lvl_pts (level 1..)
  int nlevels = 0;
  int nslow = 0;                   // points needing multi-cell traces
  // the level-ordered sweep's entries before they are grouped by source
  // class (ring_octant_base): ring level of each point, level offsets
  // [ring_nlev+1], per position the point (>= 0) or -1-b for boundary point
  // b, its start cell, and the position of each point; empty for octants
  // without ordinates
  std::vector<int32_t> ring_level, ring_off, ring_code, ring_cell, ring_pos;
  int ring_nlev = 0;
};

struct Plan {
  at3d_scene_desc d;               // copy of scalar parameters (pointers cleared)
  // sizes
  int nx, ny, nz, nx1, ny1, nxc, nyc, bcflag, ipflag, npts, ncells;
  int nmu, nphi, nphi0max, nang, ml, mm, nlm, nleg, nlegp, numphase, npart, nscatangle;
  // grid
  std::vector<float> xgrid, ygrid, zgrid, zlevels, gridpos;
  std::vector<int32_t> gridptr, neighptr;
  std::vector<int8_t> cellflags;
  // adaptive cell splitting: the cell tree (empty on a base grid), the base
  // grid's sizes and the array capacities the reference's memory parameters
  // give (MAXIG points, MAXIC cells, MAXIV source terms, MAXIDO = MAXIG *
  // NPHI0MAX), which bound the splitting as in the reference
  std::vector<int32_t> treeptr;    // [ncells][2] parent, first child (-1: none)
  int nbpts = 0, nbcells = 0;
  int maxig = 0, maxic = 0, maxiv = 0, maxido = 0;
  // property grid (kept for the direct beam and the render phase table)
  std::vector<float> extinctp, albedop, phasewtp, legenp;
  std::vector<int32_t> iphasep;
  // angles
  std::vector<float> mu, wtmu, phi, wtdo;
  std::vector<int32_t> nphi0;
  std::vector<Ordinate> ords;
  std::vector<int32_t> lofj, mofj;
  std::vector<float> ymat, wmat, ylmsun;   // [nang][nlm], [nang][nlm], [nlm]
  // optics on the RTE grid
  std::vector<float> extinct, albedo, total_ext;  // [npart][npts], [npts]
  std::vector<int32_t> iphase;                    // [npart][npts][npw]
  std::vector<float> phaseinterpwt;               // [npart][npts][npw]
  std::vector<float> legen;                       // [numphase][nleg+1] (post delta-M prep)
  std::vector<float> chi, alb, tc;                // [npts][ml+1], [npts], [npts][ml+1]
  float albmax = 0.f;
  // the per-point medium (optics, source and single-scatter constants,
  // two-stream start) is held here only when host_medium; a plan built for
  // the device set-up leaves it to the state and makes it here on request
  // (ensure_host_medium)
  bool medium_on_device = false, host_medium = true;
  double extmin = 0, scatmin = 0;
  // polarization (NSTOKES = 3): the six-element input table [6][numphase]
  // [nlegp+1] (element 1 is also legenp), the prepared elements 2, 3 and 5
  // (a2, a3, b1) like legen, per point their albedo-weighted delta-M terms
  // [npts][3][ml+1] and the b1 single-scatter terms [npts][ml+1], the I-Q
  // phase table [numphase][nscatangle], and the six real generalised
  // spherical harmonic tables of the ordinates for the transforms [6][nang][nlm]
  int nst = 1;
  std::vector<float> legenp6, legen2, legen3, legen5, chip, tms5, phasetab2, ypol, wpol;
  // the prepared six-element table in the reference's LEGEN(6, 0:NLEG,
  // NUMPHASE) order ([numphase][nleg+1][6]), for RADIANCE_TRUNCATION's
  // flat indexing
  std::vector<float> legen6flat;
  // phase tables per property point (num_micro, MAXNMICRO) and phase pairs
  // per grid point and species (8 x nmicro)
  int nmicro = 1, npw = 8;
  // INTERPMETHOD 'OO': one table per point and species, the corner's with
  // the most scattering
  bool interp_o = false;
  // rendering constants
  std::vector<float> phasetab;                    // [numphase][nscatangle]
  std::string phasetab_err;                       // negativity failure, raised by render
  std::vector<float> tms;                         // [npts][ml+1] sum_species DA * chi'_l per unit direct flux
  std::vector<int32_t> ss_off, ss_iph;            // CSR per point of (phase index, coefficient)
  std::vector<float> ss_coef;
  // solution setup
  std::vector<float> radiance0;                   // [npts][4]
  std::vector<int32_t> bctop, bcbot;
  // per-octant sweep order and levels: geometry only (grid and ordinate
  // directions), shared by plans of the same geometry (see build_levels)
  std::shared_ptr<const std::array<OctantPlan, 8>> oct_store;
  const OctantPlan* oct = nullptr;   // oct_store's eight octants
  int nlevels_max = 0;
};

// Builds the plan; returns AT3D_OK or an error status with a message.
int build_plan(const at3d_scene_desc& desc, Plan& p, std::string& err);

// The plan's per-point medium on the host, made once when the plan was built
// for the device set-up (thread safe; the plan object is filled in place).
int ensure_host_medium(const Plan& p, std::string& err);

// After the grid of a plan changed (cells split, points appended, the
// per-species optics of every point in place): the per-point source and
// render constants, boundary point lists and the sweep order and levels of
// the new grid.
int plan_refresh_points(Plan& p, std::string& err);

// A refined grid given by the caller (a saved solution's): checked for range
// and tree consistency, then the optics of every point from the property
// grid and the derived plan data as plan_refresh_points.  p is a copy of the
// base plan; pointers are 0-based as in GridTopo.
int plan_set_grid(Plan& p, int npts, int ncells, const float* gridpos, const int32_t* gridptr,
                  const int32_t* neighptr, const int32_t* treeptr, const int8_t* cellflags, std::string& err);

// The property grid and the prepared Legendre table of a plan, as the
// shared per-point set-up (medium_point.h) reads them.
PropGrid plan_prop_grid(const Plan& p);
LegTable plan_leg_table(const Plan& p);
// The sky radiance's flux weight over the downwelling ordinates (INIT_SOLUTION,
// shdomsub1.f:355-361), the two-stream start's top boundary.
float sky_flux_weight(const Plan& p);

// TRILIN_INTERP_PROP for species ipa at one point, before delta-M scaling:
// extinction, albedo and the eight (table index, weight) pairs.
void point_optics(const Plan& p, int ipa, float X, float Y, float Z, float& ext, float& alb, int32_t* iph,
                  float* wt);

// DIRECT_BEAM_PROP at one point (host trace); false if the ray leaves the
// property grid.
bool direct_beam_point(const DirectBeamGeom& b, float xi, float yi, float zi, float& dirflux,
                       std::vector<std::pair<int32_t, float>>* path_rec);

// Real generalised spherical harmonics of one direction (YLMALL_UNPOL,
// shdomsub2.f:4490-4539), J-ordered, nlm values.
void ylm_direction(float mu, float phi, int ml, int mm, float* yr);
// The six-element set of one direction (YLMALL with NSTLEG = 6,
// shdomsub2.f:4244-4359), [nlm][6].
void ylm_direction_pol(float mu, float phi, int ml, int mm, float* yr6);
// Wigner d^n_{m,m1}(x), n = 0..nrank (WIGNERFCT, shdomsub2.f:4604-4646).
void wigner_d(double x, int nrank, int m, int m1, std::vector<double>& d);

// Mixed phase function of species ipa at point ip (the reference's LEGENT,
// 'N' interpolation), coefficients 0..nleg.
void plan_mixed_legen(const Plan& p, int ipa, int ip, std::vector<float>& out);
// Source class of each point: the highest Legendre order with a nonzero
// albedo-weighted phase term; the point's source has no terms above it.
void plan_source_class(const Plan& p, std::vector<int32_t>& cls);

// Per RTE point, the property-grid points on its path to the sun and the
// derivative of the slant optical path with respect to their (delta-M scaled)
// extinction, as CSR (off[npts+1], 0-based ptr, val).
// The direct beam's trace geometry and delta-M scaled extinction
// (DIRECT_BEAM_PROP's set-up); b.zl points into p, b.ext into extdirp.
void direct_beam_setup(const Plan& p, std::vector<float>& extdirp, DirectBeamGeom& b);
// The same geometry without the extinction (b.ext null; b.zl into p).
void direct_beam_geometry(const Plan& p, DirectBeamGeom& b);
// The direct flux of every grid point on the host threads (a solve makes
// its own on the device; the plan holds none).
// Fills op's ring_* entry order for octant o from its levels (geometry only,
// so it is built once with the levels and shared like them).
bool ring_octant_base(const Plan& p, int o, OctantPlan& op, std::string& err);
int plan_direct_host(const Plan& p, std::vector<float>& dirflux, std::string& err);
int plan_sun_paths(const Plan& p, std::vector<int32_t>& off, std::vector<int32_t>& ptr,
                   std::vector<float>& val, std::string& err);

}  // namespace at3d
