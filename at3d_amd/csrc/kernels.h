// Host-side launch wrappers of the gfx950 kernels (one translation unit each).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "medium_point.h"
#include "shdom_geom.h"

namespace at3d {

// Device control of a solve's iterations, so that they run from captured
// HIP graphs without the host between them (device memory, one per state).
// Every iteration kernel returns at once while `stop` is set: k_iter_begin
// sets it when the solve has converged or reached maxiter, k_solcrit when a
// kernel flagged an error.
struct IterCtl {
  int stop;         // nonzero: the rest of this solve call's kernels do nothing
  int iters;        // iterations run in this solve call
  int total;        // iterations of the solution (state's count at the call + iters)
  int fixsh;        // FIXSH for the source pass (sticky within a call)
  float solcrit;    // SOLCRIT after the last iteration
  float a;          // the acceleration parameter of the last iteration (0: none)
  float deljdot, deljold, deljnew, jnorm;
  // parameters of the call
  int maxiter;      // total iteration limit
  int iterfixsh;
  int accel;        // acceleration on
  int one_iter;     // ALBMAX < SOLACC: SOLCRIT is set to SOLACC after an iteration
  float solacc;
  int pad[3];
};

// ---- transforms.hip: FP32 MFMA SH <-> DO transforms ----------------------
// dosrc[c][p] = max(0, sum_J S[p][J] * Yc[c][J]) * ext[p]
// S [mp][ld], Yc [ncp][ld], dosrc [ncols][ldp] (ordinate-major, ldp = mp);
// mp, ncp multiples of 64; ld multiple of 16; ext padded to mp.
void launch_sh_to_do(hipStream_t st, const float* S, const float* Yc, const float* ext,
                     float* dosrc, int mp, int ncp, int ld, int ldp, int npts, int nang,
                     const int* stop = nullptr);
// Level-ordered variant: dosrc[c][q] = max(0, sum_J S[perm_o[q]][J] * Yc[c][J]) * ext[perm_o[q]]
// for each column c of octant o, perm_o = oct_perm + o * mp (the octant's
// positions: its levels, boundary points included).  tiles[t] = {octant,
// first column, columns (<= 48), 0}.  S[p][J] must be zero for J >= ns[p].
void launch_sh_to_do_perm(hipStream_t st, const float* S, const float* Yc, const float* ext,
                          const int* ns, const int32_t* oct_perm, const int4* tiles, int ntiles,
                          float* dosrc, int mp, int ld, int ldp, int npts, const int* stop = nullptr);
// The polarized Q and U rows of the point-ordered SH->DO (no clamp): S the
// Q plane, the U plane `plane` floats on; Yc [ncp][2 ld] with ncp = 2 ldo
// rows, the Q rows then the U rows; dosrc the Q plane (U rows follow).
void launch_sh_to_do_qu(hipStream_t st, const float* S, size_t plane, const float* Yc, const float* ext,
                        float* dosrc, int mp, int ncp, int ld, int ldp, int npts, const int* stop = nullptr);
// The polarized Q and U rows of the same (NSTOKES = 3): S holds the Q plane
// and, `plane` floats after it, the U plane ([mp][ld] each); Yc rows are
// 2 ld long (the Q-plane terms then the U-plane terms), one row per output
// column of a Q / U tile; no clamp at zero.
void launch_sh_to_do_perm_qu(hipStream_t st, const float* S, size_t plane, const float* Yc, const float* ext,
                             const int* ns, const int32_t* oct_perm, const int4* tiles, int ntiles,
                             float* dosrc, int mp, int ld, int ldp, int npts, const int* stop = nullptr);
// R[p][J] = (J < nr[p]) ? sum_c dorad[c][p] * Wc[c][J] : 0
// dorad [ldo][ldp] (rows beyond nang zero), Wc [ldo][ld] (rows beyond nang zero), R [mp][ld].
void launch_do_to_sh(hipStream_t st, const float* dorad, const float* Wc, const int* nr, float* R,
                     int mp, int ld, int ldo, int ldp, int npts, const int* stop = nullptr);
// The polarized Q and U planes of R (R_Q at R, R_U `plane` floats after it)
// from the Q and U rows of dorad (2 ldo rows from `dorad`): Wc [2 ldo][2 ld],
// columns 0..ld-1 into R_Q and ld..2ld-1 into R_U, the same truncation mask.
void launch_do_to_sh_qu(hipStream_t st, const float* dorad, const float* Wc, const int* nr, float* R, size_t plane,
                        int mp, int ld, int ldo, int ldp, int npts, const int* stop = nullptr);

// ---- sweep.hip: discrete-ordinate sweeps --------------------------------
struct SweepArgs {
  GridTopo g;
  const OrdGeom* ordg;       // [ncols] device column order
  const int8_t* col_oct;     // [ncols] octant of each column
  const float* skyrad;       // [ncols]
  const float* bcrad_bot;    // [nbot]
  const int32_t* bctop;      // [ntop]
  const int32_t* bcbot;      // [nbot]
  int ntop, nbot;
  const int8_t* preset;      // [npts] bit0 top point, bit1 bottom point
  const uint8_t* face_ok;    // [8][npts] bit d: the start cell's far face across dimension d (x, y, z) has
                             // every corner computed earlier (preset or lower rank) within the ring window
  const int32_t* rank;       // [8][npts]
  const int32_t* level;      // [8][npts]
  const int32_t* lvl_pts;    // concatenated over octants (~pt = general trace)
  const int32_t* lvl_cell;
  const int32_t* lvl_off;    // per octant o: level L spans lvl_pts[lvl_off[lo_base[o]+L-1] ..
                             // lvl_off[lo_base[o]+L])
  int lo_base[8], nlev[8];
  const float* ext;          // [npts]
  const float* srcext;       // [ncols][ldp]
  float* rad;                // [ncols][ldp]
  const float* xgrid;
  const float* ygrid;
  const float* zgrid;
  int nx1, ny1, nz;
  int ldp, npts;
  int force_general;
  int* errflag;
  // level-ordered sweep (k_sweep_ring): srcext rows are in each octant's
  // ring order (levels, boundary points as entries of their own), radiances
  // stay point-ordered.  Per octant o, arrays of npts entries at o * npts.
  const int32_t* oct_pos;    // [8][npts] point -> position
  const int32_t* ring_code;  // [8][npts] position -> point, or -1-b for boundary point b
  const int32_t* ring_cell;  // [8][npts] start cell of a point entry
  const int32_t* ring_level; // [8][npts] ring level of each point
  const int8_t* ring_staged; // [8][npts] point written by its sector
  const int32_t* ring_off;   // [8][gld] level L spans positions ring_off[o*gld+L-1] .. [o*gld+L]
  int ring_nlev[8];
  int oct_col0[8], oct_nord[8];  // device columns of each octant
  const int32_t* sec_off;    // [8][gld+1] sectors completing at level L: [o*(gld+1)+L] .. [+L+1]
  const int32_t* sec_p0;     // sectors of all octants, octant o from sec_base[o]
  const uint4* sec_slot;     // eight 16-bit ring slots per sector (four used with 4-point sectors)
  int sec_base[8];
  const int32_t* top_index;  // [npts] index in bctop or -1
  const int32_t* bot_index;  // [npts] index in bcbot or -1
  const uint4* rec;          // [ncols][npts] step records (SweepRec)
  const float4* geo;         // [ncols][ncls] one-cell step of each class: u, v, path, face
  const int2* geo_rep;       // [8][ncls] a (point, start cell) of each class per octant
  const float* ext_pos;      // [8][ldp] extinction in each octant's ring order
  int ncls;                  // step classes: nz * 4 (point level x the start cell's IP flags)
  unsigned long long nz_magic;  // ceil(2^40 / nz): point level = pt - nz * ((pt * nz_magic) >> 40)
  float cls_tol;             // u, v and relative path-length tolerance of a step against its class
  const uint8_t* gflag;      // [ncols][gld] level L holds a multi-cell trace
  int* col_general;          // [ncols] nonzero: the column has multi-cell traces
  int gld;
  int ring_full;             // every column in the full instantiation
  int slots;                 // ring slots
  int sector_points;         // 8 or 4
  int* ngeneral;             // multi-cell traces found by the record pass
  const int* stop = nullptr; // IterCtl::stop of the state, or null
  // polarization: Stokes components swept (1 or 3; the ring launch's grid y)
  // and the stride between their srcext / rad planes.  Q and U have zero
  // boundary radiances and no clamp of the cell source at zero.
  int nst = 1;
  size_t pdo = 0;
  // the ring holds radiances alone (4-byte slots) and the face sources are
  // read from the ring-ordered source row
  int ring4 = 0;
  // k_sweep_ring: columns from this one on run their Q and U workgroups only
  // (blockIdx.y >= 1); the I workgroup returns at once
  int qu_only_from = 1 << 30;
};

// The fixed part of one ring entry of one ordinate, 16 bytes: head = point |
// class flags << 27 (the start cell's IP flags) | staged << 29 | kind << 30
// (kind 0: one-cell step, 1: multi-cell trace, 2: boundary point, 3: a
// zero-length step onto the corner in r[0], whose radiance it takes).  A
// one-cell step's face values are bilinear in (u, v) over four ring slots
// r[0..3]; u, v and the path length are those of its step class (the point's
// level and the class flags: every point of a class steps through a cell of
// the same shape), and ext0 is the extinction on the face.  A boundary entry
// holds its boundary index in r[0..1] and its radiance's bits in r[2..3].
struct SweepRec {
  uint32_t head;
  uint16_t r[4];
  float ext0;
};
static_assert(sizeof(SweepRec) == 16, "SweepRec is one 16-byte word");
constexpr int RECORD_PT_BITS = 27;
constexpr int RING_THREADS = 512;
constexpr int RING_K = 5;  // entries per thread per level with records in registers
constexpr int RING_KX = 4; // entries per thread loaded together past those (full instantiation)

// Bytes of LDS k_sweep_ring needs: the step classes (16 bytes each; ncls = 0
// for the full instantiation, which reads them from device memory), the ring
// of (radiance, source) pairs (8 bytes a slot, or radiances alone, 4), gld
// level offsets, gld + 1 sector offsets and gld level flags.
inline size_t ring_lds_bytes(int slots, int gld, int ncls, int slot_bytes = 8) {
  return (size_t)ncls * 16 + (size_t)slots * slot_bytes + (size_t)gld * 9 + 4;
}
// Allows k_sweep_ring launches with lds_fast / lds_full bytes of dynamic LDS
// (lds_fast 0: the fast instantiation is not used) for the ring of pairs or
// (ring4) of radiances; nonzero on failure.
int sweep_ring_max_lds(size_t lds_fast, size_t lds_full, bool ring4 = false);
void launch_sweep_records(hipStream_t st, const SweepArgs& a);
// face_ok of every point of the octants with ordinates.
void launch_face_mask(hipStream_t st, const SweepArgs& a);
// The step class table (geo) of every column, from each class's representative.
void launch_sweep_geo(hipStream_t st, const SweepArgs& a, int ncols);
// fast / full: launch the instantiation without / with the multi-cell,
// overflow and extra-sector paths (a column runs in exactly one of them).
void launch_sweep_ring(hipStream_t st, const SweepArgs& a, int col0, int ncol, size_t lds_fast, size_t lds_full,
                       bool fast, bool full);

void launch_sweep_preset(hipStream_t st, const SweepArgs& a, bool up, int col0, int ncol);
// One workgroup per ordinate column col0 .. col0+ncol-1, all levels of its octant.
void launch_sweep_columns(hipStream_t st, const SweepArgs& a, int col0, int ncol);
// Where the surface radiances also go: the boundary entries of the step
// records of the upwelling columns col0 .. col1-1 (rec null: nowhere else).
struct BottomRecs {
  uint4* rec;  // first-word plane of the step records
  int npts, col0, col1;
  const int8_t* col_oct;
  const int32_t* oct_pos;
};
// Lambertian surface radiance from the downwelling flux.
void launch_bottom_bc(hipStream_t st, const float* rad, int ldp, const int32_t* bcbot, int nbot,
                      const int32_t* col_of_ref, const float* fluxw, int ndown,
                      const float* dirflux, float albpi, float* bcrad, const BottomRecs& br,
                      const int* stop = nullptr);
void launch_fluxes(hipStream_t st, const float* rad, int ldp, const int32_t* col_of_ref,
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
  // RADIANCE_TRUNCATION of the next iteration, from this pass's NS and the
  // radiance's first term: NR written to nr_next
  int* nr_next;
  const float* tc;    // [npts][ml+1]
  float radmin;
  int highorder;
  const IterCtl* ctl = nullptr;  // null: `fixsh` as given; else FIXSH and stop from it
  // polarization (NSTOKES = 3): R, S, DS hold the I, Q and U planes `psh`
  // floats apart, chip [npts][3][ml+1] the a2, a3 and b1 terms
  int nst = 1;
  size_t psh = 0;
  const float* chip = nullptr;
};
constexpr int SRC_PTS_PER_BLOCK = 16;
void launch_radiance_truncation(hipStream_t st, const float* R, int ld, const float* tc,
                                const int* ns, const int* lofj, int* nr, int npts, int ml, int mm,
                                int nlm, float radmin, int fixsh, int highorder);
void launch_compute_source(hipStream_t st, const SourceArgs& a, int nblocks);
// R[p][J] = R0[p][J] for J < 4, zero for 4 <= J < max(4, nr[p]) (R is zero
// beyond NR already); then nr[p] = 4.
void launch_init_radiance(hipStream_t st, float* R, const float* R0, int* nr, int ld, int npts);

// ---------------------------------------------------------------------------
// The per-point medium on the device (medium.hip)
// ---------------------------------------------------------------------------
struct MediumArgs {
  PropGrid g;                  // property grid (device pointers)
  LegTable t;                  // prepared Legendre table (device)
  const float* legenp = nullptr;   // the raw table [numphase][nlegp+1] (device)
  int nlegp = 0;
  int npart = 0, npts = 0, nz = 0, ncol = 0;
  bool deltam = false;
  float secmu0 = 0.f, solarflux = 0.f, solarmu = 0.f, gndemis = 0.f, skyflux_weight = 0.f;
  const float* gridpos = nullptr;
  const float* zgrid = nullptr;
  // species-major optics [npart][npts], phase pairs [npart][npts][8]
  float* extinct = nullptr;
  float* albedo = nullptr;
  int32_t* iphase = nullptr;
  float* wt = nullptr;
  // per point
  float* total_ext = nullptr;
  float* alb = nullptr;
  float* chi = nullptr;        // [npts][ml+1]
  float* tc = nullptr;         // [npts][ml+1]
  float* tms = nullptr;        // [npts][ml+1] per unit direct flux
  int32_t* ss_cnt = nullptr;
  int32_t* ss_off = nullptr;   // [npts+1]
  int32_t* ss_iph = nullptr;
  float* ss_coef = nullptr;
  int32_t* cls = nullptr;
  unsigned int* albmax = nullptr;
  // two-stream start: per column 4 (2 nlayer + 3) doubles and 6 nz floats
  double* work = nullptr;
  float* fwork = nullptr;
  float* radiance0 = nullptr;  // [npts][4]
  float* beam_ext = nullptr;   // [npx * npy * npz]
  int* err = nullptr;
};
// Optics, constants and the pair counts' offsets (total pairs into *total).
void launch_medium_points(hipStream_t st, const MediumArgs& m, int32_t* block_sum, int32_t* total);
// The pairs themselves (after ss_iph / ss_coef are allocated).
void launch_medium_pairs(hipStream_t st, const MediumArgs& m);
// The direct beam's extinction and the two-stream start.
void launch_medium_start(hipStream_t st, const MediumArgs& m);
void launch_reduce_dots(hipStream_t st, const double* partial, int nblocks, double* out,
                        const int* stop = nullptr);
// ctl null: S += a DS; else with ctl's parameter, when it is positive.
void launch_accelerate(hipStream_t st, float* S, const float* DS, const int* ns, const int* nos,
                       int ld, int npts, float a, const IterCtl* ctl = nullptr, int nst = 1, size_t psh = 0);
// The start of an iteration (the loop test of SOLUTION_ITERATIONS and FIXSH)
// and its end (CALC_ACCEL_SOLCRIT from the reduced dot products), one thread each.
void launch_iter_begin(hipStream_t st, IterCtl* ctl);
void launch_solcrit(hipStream_t st, IterCtl* ctl, const double* dots, const int* errflag);

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
  int nst = 1;             // Stokes components: dsrc rows hold nst blocks of `cst`
  int cst = 0;             // directions (I, Q, U); out is [nrays][nst]
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
  // CORRECTINTERPOLATE (shdomsub2.f:2562): 0 keeps the source and extinction
  // a ray carries out of one cell as its start in the next
  int correct_interp = 1;
  float* out;
  int* errflag;
};
void launch_dir_source_finish(hipStream_t st, const float* raw, int ldd, int ndir, const float* ext,
                              const int* ns, const int* lofj, const float* tms, const float* G,
                              const int32_t* ss_off, const int32_t* ss_iph, const float* ss_coef,
                              const float* ssd, int numphase, int ml, int npts, int single_scatter,
                              int tms_on, float* dsrc);
void launch_render(hipStream_t st, const RayArgs& a);
// The polarized direction source: raw rows [npts][ldd] hold I, Q, U blocks of
// `cst` columns each; G3 [3][ndir][ml+1] the truncated single-scatter
// factors of I (element 1) and of Q and U (elements 2 and 6 with the sun's
// element 1), tms / tms5 [npts][ml+1] the a1 and b1 weights, ssd / ssq / ssu
// [ndir][numphase] the phase table's I, Q and U at each direction.
void launch_dir_source_finish_pol(hipStream_t st, const float* raw, int ldd, int cst, int ndir, const float* ext,
                                  const int* ns, const int* lofj, const float* tms, const float* tms5,
                                  const float* G3, const int32_t* ss_off, const int32_t* ss_iph,
                                  const float* ss_coef, const float* ssd, const float* ssq, const float* ssu,
                                  int numphase, int ml, int npts, int single_scatter, int tms_on, float* dsrc);
void launch_render_bottom(hipStream_t st, const float* fluxes, const float* dirflux,
                          const int32_t* bcbot, int nbot, float albpi, float* bcrad);
// raw[p][d] = sum_J S[p][J] * Yd[d][J] (no clamp, no extinction)
void launch_sh_dot(hipStream_t st, const float* S, const float* Yd, float* raw, int mp, int ndp,
                   int ld, int ldd, int npts, int ndir);
// raw[p][d] = sum_J S_Q[p][J] * Yd[d][J] + S_U[p][J] * Yd[d][ld + J] for
// ndp rows of Yd (2 ld long); S_U `plane` floats after S_Q; raw row stride ldo.
void launch_sh_dot_qu(hipStream_t st, const float* S, size_t plane, const float* Yd, float* raw, int mp,
                      int ndp, int ld, int ldo, int npts);

// ---------------------------------------------------------------------------
// Levis gradient
// ---------------------------------------------------------------------------
#define AT3D_MAX_L 63   // highest Legendre order the gradient kernels hold in registers

struct AdjArgs {
  RayArgs r;                   // geometry, dsrc (source x ext), rays, boundary
  const float* ss8;            // [npts][ldd] exact single scatter x ext
  const double* adj_weight;    // [nrays] adjoint x ray x Stokes weight
  const float* dirflux;        // [npts]
  float opi_gndalbedo;         // ground albedo / pi (FL Lambertian)
  int exact_single_scatter;
  int max_intervals;           // sub-step intervals a ray may record
  double* bsrc;                // [npts][ldd] source-gradient weights
  double* erad;                // [npts] radiance-term weights
  double* beamw;               // [npts] direct-beam weights
  // Jacobian rows (GRAD_INTEGRATE_1RAY): per ray, the same three weights with
  // the ray x Stokes weight alone, kept only at the grid points whose
  // interpolation kernel (rel_s) or sun path (rel_b) reaches a Jacobian point
  const double* jac_weight;    // [nrays] or null
  const int32_t* rel_s;        // [npts] slot or -1
  const int32_t* rel_b;        // [npts] slot or -1
  int nrel_s, nrel_b;
  double* jsrc;                // [nrays][nrel_s]
  double* jrad;                // [nrays][nrel_s]
  double* jbeam;               // [nrays][nrel_b]
};

struct GradPointArgs {
  int npts, ndir, ldd, ld, ml, numder, npart, maxpg, numphase, dnumphase;
  float secmu0;
  const double* bsrc;
  const double* erad;
  const double* beamw;
  double* grad;                // [numder][maxpg]
  const float* R;              // [npts][ld] radiance SH
  const float* Yd;             // [ndir][ld]
  const int* nr;
  const int* lofj;
  const float* dirflux;
  const float* ssd;            // [ndir][numphase] phase table at each direction's scattering angle
  const float* dssd;           // [ndir][dnumphase] derivative phase table, same angles
  const int32_t* partder;      // [numder] 0-based species
  const int32_t* doexact;      // [numder]
  const int32_t* sj_off;       // [npart*npts+1] CSR of the SINGSCATJ terms
  const int32_t* sj_iph;
  const float* sj_coef;
  const float* albedo;         // [npart][npts]
  const float* srcpre;         // [npts][ldd]
  const int8_t* scatj_ok;      // [npart][npts]
  const float* legt_s;         // [npart][npts][ml+1] delta-M scaled mixed phase (npart > 1)
  const float* f_s;            // [npart][npts]
  const float* legt_u;         // [npart][npts][ml+1] unscaled mixed phase
  const float* optinterpwt;    // [npts][8]
  const int32_t* interpptr;    // [npts][8]
  const float* extinctp;       // [npart][maxpg]
  const float* albedop;
  const int32_t* iphasep;      // 1-based
  const float* phasewtp;
  const float* dext;           // [numder][maxpg]
  const float* dalb;
  const float* dphasewtp;
  const int32_t* diphasep;     // 0-based
  const float* dextm;          // [numder][maxpg]
  const float* legp;           // [npart][maxpg][ml+1] unscaled phase x weight at property points
  const float* dlegp;          // [numder][maxpg][ml+1]
  const float* dfj;            // [numder][npts][8]
  const float* dalbm;
  const int32_t* bp_off;       // [maxpg+1] sun paths by property point
  const int32_t* bp_ip;
  const float* bp_w;
};

void launch_dir_source_grad(hipStream_t st, const float* raw, int ldd, int ndir, const float* ext,
                            const int* ns, const int* lofj, const float* tms, const float* G,
                            const int32_t* ss_off, const int32_t* ss_iph, const float* ss_coef,
                            const float* ssd, int numphase, int ml, int npts, int single_scatter,
                            float* dsrc, float* srcpre, float* ss8);
struct JacArgs {
  int npix, njac, numder, ndir, maxpg, nrel_s, nrel_b, ess;
  const int32_t* ray_off;      // [npix+1] first ray of each pixel
  const int32_t* dir_of_ray;   // [nrays]
  const int32_t* jac_ib;       // [njac] property point of each Jacobian column
  const int32_t* js_off;       // [njac+1] (source slot, neighbour) pairs of each column
  const int32_t* js_r;
  const int32_t* js_nb;
  const int32_t* jb_off;       // [njac+1] (beam slot, d tau / d ext) pairs of each column
  const int32_t* jb_r;
  const float* jb_w;
  const int32_t* p_of_rs;      // [nrel_s] grid point of each source slot
  const float* g8t;            // [nrel_s][ndir][numder][8] GRAD8 terms
  const float* dextm;          // [numder][maxpg]
  const float* optinterpwt;    // [npts][8]
  const double* jsrc;
  const double* jrad;
  const double* jbeam;
  float* jacobian;             // [npix][njac][numder]
};

void launch_adjoint_rays(hipStream_t st, const AdjArgs& a);
// GRAD8 of every source slot in every ray direction, then the Jacobian.
void launch_jacobian(hipStream_t st, const GradPointArgs& g, const JacArgs& j, float* g8t);
void launch_grad_points(hipStream_t st, const GradPointArgs& a);
void launch_grad_beam(hipStream_t st, const GradPointArgs& a);

}  // namespace at3d
