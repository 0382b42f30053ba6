// Device-resident SHDOM source-function iteration for gfx950 (host orchestration).
//
// One iteration (SOLUTION_ITERATIONS loop body, shdomsub1.f:674-797):
//   radiance truncation  (RADIANCE_TRUNCATION :1615)          source.hip
//   SH -> DO transform   (SH_TO_DO_UNPOL :2928) MFMA GEMM      transforms.hip
//   downwelling sweeps, Lambertian surface, upwelling sweeps
//                        (PATH_INTEGRATION :1836, BACK_INT_GRID3D_UNPOL :3695)
//                        one workgroup per ordinate over per-solve step records,
//                        replayed from a HIP graph                               sweep.hip
//   DO -> SH transform   (DO_TO_SH_UNPOL :3164) MFMA GEMM      transforms.hip
//   source function      (COMPUTE_SOURCE :967)                 source.hip
//   criterion            (CALC_ACCEL_SOLCRIT shdom_nompi.f:317) host, single precision
//   acceleration         (ACCELERATE_SOLUTION :1807)           source.hip
//
// Device layouts: SH fields [point][ld] (ld = nlm rounded up to 64); DO fields
// ordinate-major [ldo][nptsp] with the ordinates grouped by octant (ldo = nang
// rounded up to 64, nptsp = points rounded up to 64, padding zero).  With the
// level-ordered sweep the DO source rows are in their octant's position order
// (boundary points, then sweep levels) and the DO radiance rows by point.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cfloat>
#include <string>
#include <vector>

#include "devmem.h"
#include "host_plan.h"
#include "direct_beam.h"
#include "kernels.h"
#include "ring_plan.h"
#include "shdom_geom.h"
#include "solver_state.h"
#include "split_grid.h"

namespace at3d {

#define HIPCHK(x)                                                              \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      err = std::string(#x) + ": " + hipGetErrorString(e_);                    \
      return AT3D_ERR_DEVICE;                                                  \
    }                                                                          \
  } while (0)

#define TRY(x)              \
  do {                      \
    int rc_ = (x);          \
    if (rc_) return rc_;    \
  } while (0)

static size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

// Device allocations of a state are filled on the state's own stream (which
// does not synchronise with the legacy default stream, so states driven from
// several host threads never wait on, or disturb a graph capture of, each
// other) and waited for before returning.
template <typename T>
static int upload(State& s, T*& dst, const T* src, size_t n, std::string& err) {
  const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  HIPCHK(dev_alloc((void**)&dst, bytes));
  s.allocs.emplace_back(dst, bytes);
  s.device_bytes += bytes;
  if (n) {
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s.stream));
    HIPCHK(hipStreamSynchronize(s.stream));
  }
  return AT3D_OK;
}

template <typename T>
static int alloc_only(State& s, T*& dst, size_t n, std::string& err) {
  const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  HIPCHK(dev_alloc((void**)&dst, bytes));
  s.allocs.emplace_back(dst, bytes);
  s.device_bytes += bytes;
  return AT3D_OK;
}

template <typename T>
static int alloc0(State& s, T*& dst, size_t n, std::string& err) {
  const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  HIPCHK(dev_alloc((void**)&dst, bytes));
  s.allocs.emplace_back(dst, bytes);
  s.device_bytes += bytes;
  HIPCHK(hipMemsetAsync(dst, 0, bytes, s.stream));
  HIPCHK(hipStreamSynchronize(s.stream));
  return AT3D_OK;
}

static SweepArgs sweep_args(const State& s) {
  const Plan& p = *s.plan;
  SweepArgs a;
  a.g = GridTopo{s.gridpos, s.gridptr, s.neighptr, s.cellflags, p.npts, p.ncells, s.treeptr};
  a.ordg = s.ordg;
  a.skyrad = s.skyrad;
  a.bcrad_bot = s.bcrad_bot;
  a.bctop = s.bctop;
  a.bcbot = s.bcbot;
  a.ntop = (int)p.bctop.size();
  a.nbot = (int)p.bcbot.size();
  a.preset = s.preset;
  a.rank = s.rank;
  a.level = s.level;
  a.lvl_pts = s.lvl_pts;
  a.lvl_cell = s.lvl_cell;
  a.lvl_off = s.lvl_off;
  a.col_oct = s.col_oct;
  for (int o = 0; o < 8; o++) {
    a.lo_base[o] = s.lo_base_h[o];
    a.nlev[o] = s.oct_nord[o] > 0 ? p.oct[o].nlevels : 0;
  }
  a.ext = s.ext;
  a.srcext = s.dosrc;
  a.rad = s.dorad;
  a.xgrid = s.xgrid;
  a.ygrid = s.ygrid;
  a.zgrid = s.zgrid;
  a.nx1 = p.nx1;
  a.ny1 = p.ny1;
  a.nz = p.nz;
  a.ldp = s.nptsp;
  a.npts = p.npts;
  a.force_general = s.force_general;
  a.errflag = s.errflag;
  a.face_ok = s.face_ok;
  a.oct_pos = s.oct_pos;
  a.ring_code = s.ring_code;
  a.ring_cell = s.ring_cell;
  a.ring_level = s.ring_level;
  a.ring_staged = s.ring_staged;
  a.ring_off = s.ring_off;
  a.sec_off = s.sec_off;
  a.sec_p0 = s.sec_p0;
  a.sec_slot = s.sec_slot;
  for (int o = 0; o < 8; o++) {
    a.ring_nlev[o] = s.ring_nlev[o];
    a.sec_base[o] = s.sec_base[o];
    a.oct_col0[o] = s.oct_col_base[o];
    a.oct_nord[o] = s.oct_nord[o];
  }
  a.top_index = s.top_index;
  a.bot_index = s.bot_index;
  a.rec = s.rec;
  a.geo = s.geo;
  a.geo_rep = s.geo_rep;
  a.ext_pos = s.ext_pos;
  a.ncls = s.ncls;
  a.nz_magic = s.nz_magic;
  a.cls_tol = s.cls_tol;
  a.gflag = s.gflag;
  a.col_general = s.col_general;
  a.gld = s.gld;
  a.ring_full = s.ring_full;
  a.slots = s.slots;
  a.sector_points = s.sector_points;
  a.ngeneral = s.ngeneral_d;
  a.stop = nullptr;
  a.nst = s.nst;
  a.pdo = s.pdo;
  a.ring4 = s.ring4;
  return a;
}

// The level-ordered sweep, where its rings fit, unless its levels are chains
// of a point or two (periodic grids, whose multi-cell traces depend on the
// points swept just before): there the point-ordered sweep's lighter levels
// are faster (16^3 periodic, NMU 8: 22 against 45 ms per iteration).
static bool ring_active(const State& s) {
  return s.ring_ok && s.use_ring && !s.force_general && (s.use_ring == 2 || !s.ring_chain);
}

// The level-ordered sweep: each octant's ring order (its levels, with the
// hemisphere's boundary points as entries of their own), the SH->DO column
// tiles, the ring sizes, the sectors, and the per-solve step records.  Leaves
// ring_ok false (and the point-ordered sweep in use) when the rings do not fit
// in LDS or the records do not fit in memory.
static int setup_ring(const Plan& p, State& s, std::string& err) {
  s.ring_ok = false;
  // the step classes and ring layout assume the base grid's regular cells
  if (!p.treeptr.empty()) return AT3D_OK;
  bool used[8];
  for (int o = 0; o < 8; o++) used[o] = s.oct_nord[o] > 0;
  int dev = 0, maxlds = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  // a ring of (radiance, source) pairs where it fits in LDS, else (or when
  // asked for) of radiances alone
  RingPlan rp;
  const std::vector<int32_t>* cls = p.medium_on_device ? &s.medium_cls_h : nullptr;
  s.ring4 = s.force_ring4 ? 1 : 0;
  if (!build_ring_plan(p, used, (size_t)maxlds, rp, err, cls, s.ring4 ? 4 : 8)) return AT3D_ERR_NUMERIC;
  if (!s.ring4 && ring_lds_bytes(rp.slots, rp.nlev_max + 2, 0, 8) > (size_t)maxlds) {
    s.ring4 = 1;
    rp = RingPlan();
    if (!build_ring_plan(p, used, (size_t)maxlds, rp, err, cls, 4)) return AT3D_ERR_NUMERIC;
  }
  const float* total_ext = p.medium_on_device ? s.medium_ext_h.data() : p.total_ext.data();
  s.ring_chain = false;
  for (int o = 0; o < 8; o++)
    if (used[o] && rp.oct[o].nlev > 0 && (double)p.npts / rp.oct[o].nlev < 4.0) s.ring_chain = true;
  s.sector_points = rp.sector_points;
  s.slots = rp.slots;
  s.gld = rp.nlev_max + 2;
  const size_t np = p.npts;
  s.ncls = p.nz * 4;
  std::vector<int32_t> perm((size_t)8 * s.nptsp, 0);
  std::vector<float> extpos((size_t)8 * s.nptsp, 0.f);
  // a representative (point, start cell) of each step class, per octant: the
  // class's point nearest the domain's centre column (away from the seams
  // and open edges, whose steps can be degenerate)
  std::vector<int2> reps((size_t)8 * s.ncls, make_int2(-1, -1));
  float xmid = 0.f, ymid = 0.f;
  for (size_t i = 0; i < np; i++) {
    xmid = std::max(xmid, p.gridpos[i * 3]);
    ymid = std::max(ymid, p.gridpos[i * 3 + 1]);
  }
  xmid *= 0.5f;
  ymid *= 0.5f;
  std::vector<int32_t> roff((size_t)8 * s.gld, 0), soff((size_t)8 * (s.gld + 1), 0), sp0;
  std::vector<uint16_t> sslot;
  s.oct_pos_h.assign(8 * np, 0);
  std::vector<int4> tiles;
  s.nsectors = 0;
  // per-octant slices of the device arrays, octants on their own host threads
  parallel_for(8, [&](int o) {
    int32_t* pm = perm.data() + (size_t)o * s.nptsp;
    for (int q = 0; q < s.nptsp; q++) pm[q] = q;  // unused octants and padding rows
    if (!used[o]) {
      for (size_t i = 0; i < np; i++) s.oct_pos_h[o * np + i] = (int32_t)i;
      return;
    }
    const RingOctant& ro = rp.oct[o];
    std::copy(ro.pos.begin(), ro.pos.end(), s.oct_pos_h.begin() + o * np);
    const std::vector<int32_t>& bc = ((o >> 2) & 1) ? p.bcbot : p.bctop;
    float* ex = extpos.data() + (size_t)o * s.nptsp;
    int2* rp_o = reps.data() + (size_t)o * s.ncls;
    std::vector<float> dbest(s.ncls, 3.0e38f);
    for (size_t q = 0; q < np; q++) {
      pm[q] = ro.code[q] >= 0 ? ro.code[q] : bc[-1 - ro.code[q]];
      ex[q] = total_ext[pm[q]];
      if (ro.code[q] >= 0 && ro.cell[q] >= 0) {
        const int pt = ro.code[q];
        const int cls = (pt % p.nz) * 4 + (p.cellflags[ro.cell[q]] & 3);
        const float d = fabsf(p.gridpos[pt * 3] - xmid) + fabsf(p.gridpos[pt * 3 + 1] - ymid);
        if (d < dbest[cls]) {
          dbest[cls] = d;
          rp_o[cls] = make_int2(pt, ro.cell[q]);
        }
      }
    }
  });
  for (int o = 0; o < 8; o++) {
    s.ring_nlev[o] = 0;
    s.sec_base[o] = (int)sp0.size();
    if (!used[o]) continue;
    const RingOctant& ro = rp.oct[o];
    s.ring_nlev[o] = ro.nlev;
    for (int L = 0; L <= ro.nlev; L++) roff[(size_t)o * s.gld + L] = ro.off[L];
    for (int L = 0; L <= ro.nlev + 1; L++) soff[(size_t)o * (s.gld + 1) + L] = ro.sec_off[L];
    sp0.insert(sp0.end(), ro.sec_p0.begin(), ro.sec_p0.end());
    sslot.insert(sslot.end(), ro.sec_slot.begin(), ro.sec_slot.end());
    for (int c = s.oct_col_base[o]; c < s.oct_col_base[o] + s.oct_nord[o]; c += 48)
      tiles.push_back(make_int4(o, c, std::min(48, s.oct_col_base[o] + s.oct_nord[o] - c), 0));
  }
  s.nsectors = (int)sp0.size();
  // the fast instantiation holds a level in the threads' register slots and
  // one sector per thread
  s.ring_full = 0;
  for (int o = 0; o < 8; o++) {
    if (!used[o]) continue;
    const RingOctant& ro = rp.oct[o];
    for (int L = 1; L <= ro.nlev; L++)
      if (ro.off[L] - ro.off[L - 1] > RING_K * RING_THREADS || ro.sec_off[L + 1] - ro.sec_off[L] > RING_THREADS)
        s.ring_full = 1;
  }
  // the fast instantiation keeps the step classes in LDS beside the ring;
  // where they do not fit, every column takes the full one
  const int slot_bytes = s.ring4 ? 4 : 8;
  s.ring_lds = ring_lds_bytes(s.slots, s.gld, s.ncls, slot_bytes);
  s.ring_lds_full = ring_lds_bytes(s.slots, s.gld, 0, slot_bytes);
  if (s.ring_lds > (size_t)maxlds) s.ring_full = 1;
  if (s.ring_full) s.ring_lds = 0;
  if (s.slots > 65536 || p.npts >= (1 << RECORD_PT_BITS) || p.nz >= 8192 ||
      s.ring_lds_full > (size_t)maxlds || sweep_ring_max_lds(s.ring_lds, s.ring_lds_full, s.ring4 != 0) != 0)
    return AT3D_OK;
  s.nz_magic = ((1ull << 40) + (unsigned long long)p.nz - 1) / (unsigned long long)p.nz;
  // Steps of one class differ only by the rounding of the float grid
  // positions: a few units in the last place of the largest coordinate over
  // the smallest grid spacing (at least 2e-5)
  {
    float amax = 0.f, smin = 3.0e38f;
    for (size_t i = 0; i < (size_t)np * 3; i++) amax = std::max(amax, fabsf(p.gridpos[i]));
    auto spacing = [&](const std::vector<float>& gr) {
      for (size_t i = 1; i < gr.size(); i++)
        if (gr[i] > gr[i - 1]) smin = std::min(smin, gr[i] - gr[i - 1]);
    };
    spacing(p.xgrid);
    spacing(p.ygrid);
    spacing(p.zgrid);
    s.cls_tol = 2e-5f;
    if (smin < 3.0e38f) s.cls_tol = std::max(s.cls_tol, 8.f * FLT_EPSILON * amax / smin);
  }
  const size_t nrec = (size_t)p.nang * np;
  if (dev_alloc((void**)&s.rec, nrec * sizeof(uint4)) != hipSuccess) {
    (void)hipGetLastError();
    s.rec = nullptr;
    return AT3D_OK;
  }
  s.allocs.emplace_back(s.rec, nrec * sizeof(uint4));
  s.device_bytes += nrec * sizeof(uint4);
  std::vector<int32_t> top_index(p.npts, -1);
  for (size_t b = 0; b < p.bctop.size(); b++) top_index[p.bctop[b]] = (int32_t)b;
  TRY(upload(s, s.top_index, top_index.data(), top_index.size(), err));
  TRY(upload(s, s.oct_perm, perm.data(), perm.size(), err));
  TRY(upload(s, s.oct_pos, s.oct_pos_h.data(), s.oct_pos_h.size(), err));
  // per-octant slices straight from the ring plan (zero for octants without
  // ordinates, which no column reads)
  TRY(alloc0(s, s.ring_code, 8 * np, err));
  TRY(alloc0(s, s.ring_cell, 8 * np, err));
  TRY(alloc0(s, s.ring_level, 8 * np, err));
  TRY(alloc0(s, s.ring_staged, 8 * np, err));
  for (int o = 0; o < 8; o++) {
    if (!used[o]) continue;
    const RingOctant& ro = rp.oct[o];
    HIPCHK(hipMemcpyAsync(s.ring_code + o * np, ro.code.data(), np * 4, hipMemcpyHostToDevice, s.stream));
    HIPCHK(hipMemcpyAsync(s.ring_cell + o * np, ro.cell.data(), np * 4, hipMemcpyHostToDevice, s.stream));
    HIPCHK(hipMemcpyAsync(s.ring_level + o * np, ro.level.data(), np * 4, hipMemcpyHostToDevice, s.stream));
    HIPCHK(hipMemcpyAsync(s.ring_staged + o * np, ro.staged.data(), np, hipMemcpyHostToDevice, s.stream));
  }
  HIPCHK(hipStreamSynchronize(s.stream));
  TRY(upload(s, s.ring_off, roff.data(), roff.size(), err));
  TRY(upload(s, s.sec_off, soff.data(), soff.size(), err));
  if (sp0.empty()) sp0.push_back(0);
  sslot.resize(std::max<size_t>(sslot.size(), 8), 0);
  TRY(upload(s, s.sec_p0, sp0.data(), sp0.size(), err));
  TRY(upload(s, s.sec_slot, reinterpret_cast<const uint4*>(sslot.data()), sslot.size() / 8, err));
  TRY(upload(s, s.sh_tiles, tiles.data(), tiles.size(), err));
  if (s.nst == 3) {
    // the Q and U rows of each tile's columns: Q at the same columns of the
    // Q plane, U ldo rows on
    std::vector<int4> qu(tiles);
    for (const int4& t : tiles) qu.push_back(make_int4(t.x, t.y + s.ldo, t.z, 0));
    TRY(upload(s, s.sh_tiles_qu, qu.data(), qu.size(), err));
  }
  TRY(upload(s, s.ext_pos, extpos.data(), extpos.size(), err));
  TRY(upload(s, s.geo_rep, reps.data(), reps.size(), err));
  TRY(alloc0(s, s.geo, (size_t)p.nang * s.ncls, err));
  s.n_sh_tiles = (int)tiles.size();
  TRY(alloc0(s, s.gflag, (size_t)p.nang * s.gld, err));
  TRY(alloc0(s, s.col_general, (size_t)p.nang, err));
  TRY(alloc0(s, s.ngeneral_d, (size_t)1, err));
  TRY(alloc0(s, s.face_ok, 8 * np, err));
  launch_sweep_geo(s.stream, sweep_args(s), p.nang);
  launch_face_mask(s.stream, sweep_args(s));
  launch_sweep_records(s.stream, sweep_args(s));
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(&s.ngeneral, s.ngeneral_d, sizeof(int), hipMemcpyDeviceToHost, s.stream));
  s.col_general_h.assign(p.nang, 0);
  HIPCHK(hipMemcpyAsync(s.col_general_h.data(), s.col_general, p.nang * sizeof(int32_t),
                        hipMemcpyDeviceToHost, s.stream));
  HIPCHK(hipStreamSynchronize(s.stream));
  s.ring_ok = true;
  return AT3D_OK;
}


// Device bytes a state of this plan allocates (its arrays and the sweep
// records; not the per-call render and gradient buffers), for admission by
// memory footprint before anything is allocated.
size_t plan_footprint_bytes(const Plan& p) {
  const size_t npts = p.npts, ncells = p.ncells, nang = p.nang;
  const size_t ld = round_up(p.nlm, 64), nptsp = round_up(p.npts, 64), ldo = round_up(p.nang, 64);
  const size_t ml1 = p.ml + 1;
  size_t b = 0;
  b += 2 * ldo * ld * 4;                 // Y, W
  b += 256;                              // iteration control block
  b += 3 * nptsp * ld * 4 * p.nst;       // S, R, DS (each Stokes plane)
  b += 2 * nptsp * ldo * 4 * p.nst;      // DO source and radiance
  if (p.nst == 3) b += 4 * ldo * ld * 4 * 2 + 4 * npts * ml1 * 4;   // Q / U tables, a2 a3 b1 and b1 TMS terms
  b += npts * 16;                        // two-stream radiance terms
  b += 3 * npts * ml1 * 4;               // chi, tc, tms
  b += npts * 4 * 12;                    // per-point scalars, counts, fluxes, partials
  b += npts * 12 + ncells * 57;          // grid positions, cell topology and flags
  b += 8 * npts * 4 * 4 + npts;          // sweep order: rank, level, level lists, preset
  b += nang * npts * 16;                 // sweep records
  b += nang * (size_t)p.nz * 4 * 16 + 8 * nptsp * 4 + 8 * (size_t)p.nz * 4 * 8;  // step classes, ring-order extinction
  b += 8 * nptsp * 4 + 8 * npts * (4 * 4 + 2) + 8 * (npts / 4 + 8) * 20;  // ring layout, face masks, sectors
  if (p.host_medium)
    b += p.ss_off.size() * 4 + p.ss_iph.size() * 4 + p.ss_coef.size() * 4;
  else   // the device set-up: at most 8 pairs per species and point, its work arrays
    b += npts * 4 * (1 + 2 * (size_t)p.npw * p.npart) + npts * 4 * (4 + 3 * ml1 + (2 + 2 * (size_t)p.npw) * p.npart) +
         (size_t)p.d.npx * p.d.npy * p.d.npz * 4 * (4 * (size_t)p.npart + 1) + (npts / p.nz) * (size_t)p.nz * 80;
  b += (size_t)p.d.npx * p.d.npy * p.d.npz * 4 + p.zlevels.size() * 4;   // direct-beam extinction
  return b;
}

// The per-point medium made on the device (medium.hip) for a plan built for
// it: total extinction, mixed albedo, source constants, single-scatter
// constants (CSR), the two-stream start and the direct beam's extinction;
// the source classes and the total extinction come back for the ring layout.
static int device_medium(const Plan& p, State& s, std::string& err) {
  const at3d_scene_desc& d = p.d;
  const size_t npg = (size_t)d.npx * d.npy * d.npz, np = p.npts, ml1 = p.ml + 1;
  MediumArgs m;
  float *extp = nullptr, *albp = nullptr, *wtp = nullptr, *zl = nullptr, *leg = nullptr, *legp = nullptr;
  int32_t* iphp = nullptr;
  TRY(upload(s, extp, p.extinctp.data(), p.extinctp.size(), err));
  TRY(upload(s, albp, p.albedop.data(), p.albedop.size(), err));
  TRY(upload(s, iphp, p.iphasep.data(), p.iphasep.size(), err));
  TRY(upload(s, wtp, p.phasewtp.data(), p.phasewtp.size(), err));
  TRY(upload(s, zl, p.zlevels.data(), p.zlevels.size(), err));
  TRY(upload(s, leg, p.legen.data(), p.legen.size(), err));
  TRY(upload(s, legp, p.legenp.data(), p.legenp.size(), err));
  m.g = PropGrid{extp, albp, iphp, wtp, zl, d.npx, d.npy, d.npz, d.delx, d.dely, p.extmin, p.scatmin, p.nmicro,
                 p.interp_o};
  m.t = LegTable{leg, p.nleg + 1, p.ml, p.legen.size(), p.npw};
  m.t.interp_o = p.interp_o;
  m.legenp = legp;
  m.nlegp = p.nlegp;
  m.npart = p.npart;
  m.npts = p.npts;
  m.nz = p.nz;
  m.ncol = p.npts / p.nz;
  m.deltam = d.deltam != 0;
  m.secmu0 = (float)(1.0 / fabs((double)d.solarmu));
  m.solarflux = d.solarflux;
  m.solarmu = d.solarmu;
  m.gndemis = 1.0f - d.gndalbedo;
  m.skyflux_weight = sky_flux_weight(p);
  m.gridpos = s.gridpos;
  m.zgrid = s.zgrid;
  TRY(alloc0(s, m.extinct, (size_t)p.npart * np, err));
  TRY(alloc0(s, m.albedo, (size_t)p.npart * np, err));
  TRY(alloc0(s, m.iphase, (size_t)p.npart * np * p.npw, err));
  TRY(alloc0(s, m.wt, (size_t)p.npart * np * p.npw, err));
  TRY(alloc0(s, s.ext, (size_t)s.nptsp, err));
  m.total_ext = s.ext;
  TRY(alloc0(s, s.alb, np, err));
  TRY(alloc0(s, s.chi, np * ml1, err));
  TRY(alloc0(s, s.tc, np * ml1, err));
  TRY(alloc0(s, s.tms, np * ml1, err));
  TRY(alloc0(s, s.ss_off, np + 1, err));
  m.alb = s.alb;
  m.chi = s.chi;
  m.tc = s.tc;
  m.tms = s.tms;
  m.ss_off = s.ss_off;
  int32_t *cnt = nullptr, *cls = nullptr, *bsum = nullptr, *total = nullptr;
  unsigned int* amax = nullptr;
  TRY(alloc0(s, cnt, np, err));
  TRY(alloc0(s, cls, np, err));
  TRY(alloc0(s, bsum, (np + 1023) / 1024 + 1, err));
  TRY(alloc0(s, total, (size_t)1, err));
  TRY(alloc0(s, amax, (size_t)1, err));
  m.ss_cnt = cnt;
  m.cls = cls;
  m.albmax = amax;
  launch_medium_points(s.stream, m, bsum, total);
  HIPCHK(hipGetLastError());
  int32_t npairs = 0;
  unsigned int amax_h = 0;
  HIPCHK(hipMemcpyAsync(&npairs, total, sizeof npairs, hipMemcpyDeviceToHost, s.stream));
  HIPCHK(hipMemcpyAsync(&amax_h, amax, sizeof amax_h, hipMemcpyDeviceToHost, s.stream));
  s.medium_cls_h.resize(np);
  s.medium_ext_h.resize(np);
  HIPCHK(hipMemcpyAsync(s.medium_cls_h.data(), cls, np * sizeof(int32_t), hipMemcpyDeviceToHost, s.stream));
  HIPCHK(hipMemcpyAsync(s.medium_ext_h.data(), s.ext, np * sizeof(float), hipMemcpyDeviceToHost, s.stream));
  HIPCHK(hipStreamSynchronize(s.stream));
  memcpy(&s.albmax, &amax_h, sizeof(float));
  TRY(alloc0(s, s.ss_iph, (size_t)std::max(npairs, 1), err));
  TRY(alloc0(s, s.ss_coef, (size_t)std::max(npairs, 1), err));
  m.ss_iph = s.ss_iph;
  m.ss_coef = s.ss_coef;
  launch_medium_pairs(s.stream, m);
  // the two-stream start and the direct beam's extinction
  const size_t nlayer = p.nz - 1, ncol = m.ncol;
  TRY(alloc0(s, m.work, ncol * 4 * (2 * nlayer + 3), err));
  TRY(alloc0(s, m.fwork, ncol * (3 * (size_t)p.nz + 3 * (nlayer + 1)), err));
  TRY(alloc0(s, s.R0, np * 4, err));
  TRY(alloc0(s, s.beam_ext, npg, err));
  m.radiance0 = s.R0;
  m.beam_ext = s.beam_ext;
  m.err = s.errflag;
  launch_medium_start(s.stream, m);
  HIPCHK(hipGetLastError());
  int ef = 0;
  HIPCHK(hipMemcpyAsync(&ef, s.errflag, sizeof ef, hipMemcpyDeviceToHost, s.stream));
  HIPCHK(hipStreamSynchronize(s.stream));
  if (ef) {
    err = "EDDRTF: singular two-stream system";
    return AT3D_ERR_NUMERIC;
  }
  return AT3D_OK;
}

int state_create(const Plan& p, State& s, std::string& err) {
  s.plan = &p;
  if (!s.base_plan) s.base_plan = &p;
  HIPCHK(hipGetDevice(&s.device));
  HIPCHK(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
  s.npts = p.npts;
  s.nlm = p.nlm;
  s.nang = p.nang;
  if (p.nlm > 512) {
    err = "NLM > 512 is not supported by the source kernel";
    return AT3D_ERR_ARG;
  }
  s.ld = (int)round_up(p.nlm, 64);
  s.nptsp = (int)round_up(p.npts, 64);
  s.ldo = (int)round_up(p.nang, 64);
  s.nst = p.nst;
  s.psh = (size_t)s.nptsp * s.ld;
  s.pdo = (size_t)s.nptsp * s.ldo;
  s.ndown = 0;
  for (const Ordinate& o : p.ords) s.ndown += o.down ? 1 : 0;
  for (int k = 0; k < p.nang; k++)
    if (p.ords[k].down != (k < s.ndown)) {
      err = "internal: downwelling ordinates are not first";
      return AT3D_ERR_ARG;
    }
  // device column of each ordinate: grouped by octant, reference order inside
  s.col_of_ref.assign(p.nang, 0);
  s.ref_of_col.assign(p.nang, 0);
  int col = 0;
  for (int o = 0; o < 8; o++) {
    s.oct_col_base[o] = col;
    s.oct_nord[o] = 0;
    for (int k = 0; k < p.nang; k++)
      if (p.ords[k].g.ioct == o) {
        s.col_of_ref[k] = col;
        s.ref_of_col[col] = k;
        col++;
        s.oct_nord[o]++;
      }
  }
  // transform matrices by device column, zero padded
  std::vector<float> Y((size_t)s.ldo * s.ld, 0.f), W((size_t)s.ldo * s.ld, 0.f);
  for (int k = 0; k < p.nang; k++) {
    const int c = s.col_of_ref[k];
    for (int j = 0; j < p.nlm; j++) {
      Y[(size_t)c * s.ld + j] = p.ymat[(size_t)k * p.nlm + j];
      W[(size_t)c * s.ld + j] = p.wmat[(size_t)k * p.nlm + j];
    }
  }
  TRY(upload(s, s.Y, Y.data(), Y.size(), err));
  TRY(upload(s, s.W, W.data(), W.size(), err));
  if (p.nst == 3) {
    // the Q / U transform tables by device column (make_sh_tables_pol's
    // elements): SH->DO rows [Y2 | Y5] for Q, [Y6 | Y3] for U; DO->SH rows
    // [W2 | W6] for a Q ordinate, [W5 | W3] for a U ordinate
    const size_t row = 2 * (size_t)s.ld;
    std::vector<float> Yq((size_t)2 * s.ldo * row, 0.f), Wq((size_t)2 * s.ldo * row, 0.f);
    auto el = [&](const std::vector<float>& t, int e, int k, int j) {
      return t[((size_t)e * p.nang + k) * p.nlm + j];
    };
    for (int k = 0; k < p.nang; k++) {
      const int c = s.col_of_ref[k];
      float* yq = &Yq[(size_t)c * row];
      float* yu = &Yq[(size_t)(s.ldo + c) * row];
      float* wq = &Wq[(size_t)c * row];
      float* wu = &Wq[(size_t)(s.ldo + c) * row];
      for (int j = 0; j < p.nlm; j++) {
        yq[j] = el(p.ypol, 1, k, j);
        yq[s.ld + j] = el(p.ypol, 4, k, j);
        yu[j] = el(p.ypol, 5, k, j);
        yu[s.ld + j] = el(p.ypol, 2, k, j);
        wq[j] = el(p.wpol, 1, k, j);
        wq[s.ld + j] = el(p.wpol, 5, k, j);
        wu[j] = el(p.wpol, 4, k, j);
        wu[s.ld + j] = el(p.wpol, 2, k, j);
      }
    }
    TRY(upload(s, s.Yqu, Yq.data(), Yq.size(), err));
    TRY(upload(s, s.Wqu, Wq.data(), Wq.size(), err));
    TRY(upload(s, s.chip, p.chip.data(), p.chip.size(), err));
    TRY(upload(s, s.tms5, p.tms5.data(), p.tms5.size(), err));
  }
  std::vector<float> ysun(s.ld, 0.f);
  std::copy(p.ylmsun.begin(), p.ylmsun.end(), ysun.begin());
  TRY(upload(s, s.ysun, ysun.data(), ysun.size(), err));
  TRY(upload(s, s.lofj, p.lofj.data(), p.lofj.size(), err));
  if (!p.medium_on_device) {
    std::vector<float> ext(s.nptsp, 0.f);
    std::copy(p.total_ext.begin(), p.total_ext.end(), ext.begin());
    TRY(upload(s, s.ext, ext.data(), ext.size(), err));
    TRY(upload(s, s.alb, p.alb.data(), p.alb.size(), err));
    TRY(upload(s, s.chi, p.chi.data(), p.chi.size(), err));
    TRY(upload(s, s.tc, p.tc.data(), p.tc.size(), err));
    s.albmax = p.albmax;
  }
  TRY(alloc0(s, s.dirflux, (size_t)p.npts, err));   // made by the direct-beam pass below
  TRY(alloc0(s, s.f0, (size_t)p.npts, err));
  const size_t nsh = (size_t)s.nptsp * s.ld * s.nst, ndo = (size_t)s.nptsp * s.ldo * s.nst;
  TRY(alloc0(s, s.S, nsh, err));
  TRY(alloc0(s, s.R, nsh, err));
  TRY(alloc0(s, s.DS, nsh, err));
  TRY(alloc0(s, s.dosrc, ndo, err));
  TRY(alloc0(s, s.dorad, ndo, err));
  TRY(alloc0(s, s.ns, (size_t)p.npts, err));
  TRY(alloc0(s, s.nr, (size_t)p.npts, err));
  TRY(alloc0(s, s.nr_next, (size_t)p.npts, err));
  TRY(alloc0(s, s.nos, (size_t)p.npts, err));
  TRY(alloc0(s, s.fluxes, (size_t)2 * p.npts, err));
  s.nsrc_blocks = (p.npts + SRC_PTS_PER_BLOCK - 1) / SRC_PTS_PER_BLOCK;
  TRY(alloc0(s, s.partial, (size_t)s.nsrc_blocks * 4, err));
  TRY(alloc0(s, s.dots, (size_t)4, err));
  TRY(alloc0(s, s.errflag, (size_t)1, err));
  // sweep geometry
  TRY(upload(s, s.gridpos, p.gridpos.data(), p.gridpos.size(), err));
  TRY(upload(s, s.zgrid, p.zgrid.data(), p.zgrid.size(), err));
  if (p.medium_on_device) TRY(device_medium(p, s, err));
  {
    // MAKE_DIRECT on the device: the direct flux of every grid point and F0
    std::vector<float> extdirp;
    DirectBeamGeom beam;
    if (p.medium_on_device)
      direct_beam_geometry(p, beam);
    else
      direct_beam_setup(p, extdirp, beam);
    // a sun within 1e-6 of the horizon never reaches the top (the trace runs
    // out of steps): the reference's error, without the wait
    if (beam.cz == 0.0) {
      err = "DIRECT_BEAM_PROP: ray left the property grid";
      return AT3D_ERR_NUMERIC;
    }
    // a trace crosses each z level once and each x / y cell face of its
    // horizontal run (height / cz * |c| over the spacing) once
    const double zspan = (double)p.zlevels[p.d.npz - 1] - (double)p.zlevels[0];
    const double run = zspan / beam.cz;
    const double steps = p.d.npz + 8.0 + run * (fabs(beam.cx) / beam.delxd + fabs(beam.cy) / beam.delyd);
    beam.max_steps = (long long)std::min(1.0e8, 2.0 * steps + 64.0);
    float* d_ext = s.beam_ext;
    float* d_zl = nullptr;
    if (!p.medium_on_device) TRY(upload(s, d_ext, extdirp.data(), extdirp.size(), err));
    TRY(upload(s, d_zl, p.zlevels.data(), p.zlevels.size(), err));
    beam.ext = d_ext;
    beam.zl = d_zl;
    launch_direct_beam(s.stream, beam, s.gridpos, p.npts, 1.0f / fabsf(p.d.solarmu), s.dirflux, s.f0, s.errflag);
    HIPCHK(hipGetLastError());
    int ef = 0;
    HIPCHK(hipMemcpyAsync(&ef, s.errflag, sizeof(int), hipMemcpyDeviceToHost, s.stream));
    HIPCHK(hipStreamSynchronize(s.stream));
    if (ef) {
      err = "DIRECT_BEAM_PROP: ray left the property grid";
      return AT3D_ERR_NUMERIC;
    }
  }
  TRY(upload(s, s.gridptr, p.gridptr.data(), p.gridptr.size(), err));
  TRY(upload(s, s.neighptr, p.neighptr.data(), p.neighptr.size(), err));
  if (!p.treeptr.empty()) TRY(upload(s, s.treeptr, p.treeptr.data(), p.treeptr.size(), err));
  TRY(upload(s, s.cellflags, p.cellflags.data(), p.cellflags.size(), err));
  TRY(upload(s, s.xgrid, p.xgrid.data(), p.xgrid.size(), err));
  TRY(upload(s, s.ygrid, p.ygrid.data(), p.ygrid.size(), err));
  std::vector<OrdGeom> og(p.nang);
  std::vector<float> sky(p.nang, 0.f), fluxw(p.nang);
  for (int k = 0; k < p.nang; k++) {
    const int c = s.col_of_ref[k];
    og[c] = p.ords[k].g;
    sky[c] = p.ords[k].down ? p.d.skyrad : 0.f;
    fluxw[k] = fabsf(p.ords[k].mu) * p.ords[k].wtdo;  // reference order
  }
  TRY(upload(s, s.ordg, og.data(), og.size(), err));
  TRY(upload(s, s.skyrad, sky.data(), sky.size(), err));
  TRY(upload(s, s.fluxw, fluxw.data(), fluxw.size(), err));
  TRY(upload(s, s.col_of_ref_d, s.col_of_ref.data(), s.col_of_ref.size(), err));
  TRY(upload(s, s.bctop, p.bctop.data(), p.bctop.size(), err));
  TRY(upload(s, s.bcbot, p.bcbot.data(), p.bcbot.size(), err));
  TRY(alloc0(s, s.bcrad_bot, p.bcbot.size(), err));
  std::vector<int8_t> preset(p.npts, 0);
  for (int i : p.bctop) preset[i] |= 1;
  for (int i : p.bcbot) preset[i] |= 2;
  TRY(upload(s, s.preset, preset.data(), preset.size(), err));
  // rank and level of every octant (the records and the multi-cell traces
  // read them); the point-ordered sweep's level lists wait for its first use
  TRY(alloc_only(s, s.rank, (size_t)8 * p.npts, err));
  TRY(alloc_only(s, s.level, (size_t)8 * p.npts, err));
  s.nlevels_max = 0;
  {
    int base = 0, lob = 0;
    for (int o = 0; o < 8; o++) {
      const OctantPlan& op = p.oct[o];
      HIPCHK(hipMemcpyAsync(s.rank + (size_t)o * p.npts, op.rank.data(), (size_t)p.npts * 4,
                            hipMemcpyHostToDevice, s.stream));
      HIPCHK(hipMemcpyAsync(s.level + (size_t)o * p.npts, op.level.data(), (size_t)p.npts * 4,
                            hipMemcpyHostToDevice, s.stream));
      s.oct_pts_base_h[o] = base;
      s.lo_base_h[o] = lob;
      base += (int)op.lvl_pts.size();
      lob += op.nlevels + 1;
      if (s.oct_nord[o] > 0) s.nlevels_max = std::max(s.nlevels_max, op.nlevels);
    }
    HIPCHK(hipStreamSynchronize(s.stream));
  }
  std::vector<int8_t> coct(s.ldo, 0);
  for (int o = 0; o < 8; o++)
    for (int c = s.oct_col_base[o]; c < s.oct_col_base[o] + s.oct_nord[o]; c++) coct[c] = (int8_t)o;
  TRY(upload(s, s.col_oct, coct.data(), coct.size(), err));
  // rendering constants
  if (!p.medium_on_device) {
    TRY(upload(s, s.tms, p.tms.data(), p.tms.size(), err));
    TRY(upload(s, s.ss_off, p.ss_off.data(), p.ss_off.size(), err));
    TRY(upload(s, s.ss_iph, p.ss_iph.data(), p.ss_iph.size(), err));
    TRY(upload(s, s.ss_coef, p.ss_coef.data(), p.ss_coef.size(), err));
  }
  // the plan holds them per unit direct flux
  launch_direct_scale(s.stream, s.dirflux, p.npts, p.ml + 1, s.tms, s.ss_off, s.ss_coef);
  if (p.nst == 3) launch_direct_scale(s.stream, s.dirflux, p.npts, p.ml + 1, s.tms5, s.ss_off, nullptr);
  HIPCHK(hipGetLastError());
  TRY(alloc0(s, s.bcrad_render, p.bcbot.size(), err));
  std::vector<int32_t> bot_index(p.npts, -1);
  for (size_t b = 0; b < p.bcbot.size(); b++) bot_index[p.bcbot[b]] = (int32_t)b;
  TRY(upload(s, s.bot_index, bot_index.data(), bot_index.size(), err));
  if (!p.medium_on_device) TRY(upload(s, s.R0, p.radiance0.data(), p.radiance0.size(), err));
  if (hipHostMalloc((void**)&s.host_dots, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    s.host_dots = nullptr;
    err = "hipHostMalloc of the iteration readback buffer failed";
    return AT3D_ERR_DEVICE;
  }
  if (hipHostMalloc((void**)&s.ctl_h, sizeof(IterCtl), hipHostMallocDefault) != hipSuccess) {
    s.ctl_h = nullptr;
    err = "hipHostMalloc of the iteration control block failed";
    return AT3D_ERR_DEVICE;
  }
  TRY(alloc0(s, s.ctl, 1, err));
  s.nr_buf0 = s.nr;
  TRY(setup_ring(p, s, err));
  for (int e = 0; e < State::NEV; e++) HIPCHK(hipEventCreate(&s.ev[e]));
  for (int e = 0; e < 4; e++) HIPCHK(hipEventCreate(&s.ev_sweep[e]));
  for (int e = 0; e < 2; e++) HIPCHK(hipEventCreate(&s.ev_ray[e]));
  return AT3D_OK;
}

void state_free(State& s) {
  // the stream's work is done before its blocks go back for reuse
  if (s.stream) (void)hipStreamSynchronize(s.stream);
  {
    std::lock_guard<std::mutex> lock(device_mutex());
    if (s.sweep_exec) (void)hipGraphExecDestroy(s.sweep_exec);
    if (s.sweep_graph) (void)hipGraphDestroy(s.sweep_graph);
  }
  s.sweep_exec = nullptr;
  s.sweep_graph = nullptr;
  for (auto& a : s.allocs) dev_free(a.first, a.second);
  s.allocs.clear();
  dev_free(s.rbuf, s.rbuf_bytes);
  s.rbuf = nullptr;
  s.rbuf_bytes = 0;
  dev_free(s.gbuf, s.gbuf_bytes);
  s.gbuf = nullptr;
  s.gbuf_bytes = 0;
  for (int e = 0; e < State::NEV; e++)
    if (s.ev[e]) (void)hipEventDestroy(s.ev[e]);
  for (int e = 0; e < 4; e++)
    if (s.ev_sweep[e]) (void)hipEventDestroy(s.ev_sweep[e]);
  for (int e = 0; e < 2; e++)
    if (s.ev_ray[e]) (void)hipEventDestroy(s.ev_ray[e]);
  if (s.host_dots) (void)hipHostFree(s.host_dots);
  s.host_dots = nullptr;
  if (s.ctl_h) (void)hipHostFree(s.ctl_h);
  s.ctl_h = nullptr;
  for (int k = 0; k < 2; k++) {
    if (s.iter_exec[k]) (void)hipGraphExecDestroy(s.iter_exec[k]);
    if (s.iter_graph[k]) (void)hipGraphDestroy(s.iter_graph[k]);
    s.iter_exec[k] = nullptr;
    s.iter_graph[k] = nullptr;
  }
  if (s.stream) (void)hipStreamDestroy(s.stream);
  s.stream = nullptr;
}

int state_set_ring4(State& s, bool on, std::string& err) {
  if (s.force_ring4 == (on ? 1 : 0)) return AT3D_OK;
  HIPCHK(hipStreamSynchronize(s.stream));
  {
    std::lock_guard<std::mutex> lock(device_mutex());
    if (s.sweep_exec) (void)hipGraphExecDestroy(s.sweep_exec);
    if (s.sweep_graph) (void)hipGraphDestroy(s.sweep_graph);
    for (int k = 0; k < 2; k++) {
      if (s.iter_exec[k]) (void)hipGraphExecDestroy(s.iter_exec[k]);
      if (s.iter_graph[k]) (void)hipGraphDestroy(s.iter_graph[k]);
      s.iter_exec[k] = nullptr;
      s.iter_graph[k] = nullptr;
    }
  }
  s.sweep_exec = nullptr;
  s.sweep_graph = nullptr;
  s.force_ring4 = on ? 1 : 0;
  // the new layout's arrays are allocated afresh (the old ones are held
  // until the state is freed)
  return setup_ring(*s.plan, s, err);
}

// Replaces s by a new state of the refined grid `owned` (null: s's base
// plan), keeping the options and the iteration state; the solution arrays
// come from `f` (the refined grid's points; null: none, as before an
// initialisation).
static int state_replace(State& s, std::shared_ptr<Plan> owned, SplitFields* f, std::string& err) {
  const Plan& q = owned ? *owned : *s.base_plan;
  State n;
  n.base_plan = s.base_plan;
  n.use_ring = s.use_ring;
  n.force_general = s.force_general;
  n.force_ring4 = s.force_ring4;
  n.use_graph = s.use_graph;
  n.use_iter_graph = s.use_iter_graph;
  int rc = state_create(q, n, err);
  if (rc) {
    state_free(n);
    return rc;
  }
  if (f) {
    const size_t np = q.npts, ld = n.ld;
    if ((int)ld != f->ld || f->S.size() < np * ld || f->ns.size() < np || f->nst != n.nst ||
        (n.nst == 3 && (f->Squ[1].size() < np * ld || f->Rqu[1].size() < np * ld || f->DSqu[1].size() < np * ld))) {
      state_free(n);
      err = "internal: refined grid fields do not match the new state";
      return AT3D_ERR_ARG;
    }
    std::vector<float> flux((size_t)2 * np, 0.f);
    std::vector<float> old((size_t)2 * s.npts);
    HIPCHK(hipMemcpyAsync(old.data(), s.fluxes, old.size() * sizeof(float), hipMemcpyDeviceToHost, s.stream));
    HIPCHK(hipStreamSynchronize(s.stream));
    for (int i = 0; i < s.npts; i++) {
      flux[i] = old[i];
      flux[np + i] = old[(size_t)s.npts + i];
    }
    HIPCHK(hipMemcpyAsync(n.S, f->S.data(), np * ld * sizeof(float), hipMemcpyHostToDevice, n.stream));
    HIPCHK(hipMemcpyAsync(n.R, f->R.data(), np * ld * sizeof(float), hipMemcpyHostToDevice, n.stream));
    HIPCHK(hipMemcpyAsync(n.DS, f->DS.data(), np * ld * sizeof(float), hipMemcpyHostToDevice, n.stream));
    HIPCHK(hipMemcpyAsync(n.ns, f->ns.data(), np * sizeof(int), hipMemcpyHostToDevice, n.stream));
    HIPCHK(hipMemcpyAsync(n.nr, f->nr.data(), np * sizeof(int), hipMemcpyHostToDevice, n.stream));
    HIPCHK(hipMemcpyAsync(n.nos, f->nos.data(), np * sizeof(int), hipMemcpyHostToDevice, n.stream));
    if (n.nst == 3)
      for (int k = 0; k < 2; k++) {
        const size_t off = (size_t)(k + 1) * n.psh;
        HIPCHK(hipMemcpyAsync(n.S + off, f->Squ[k].data(), np * ld * sizeof(float), hipMemcpyHostToDevice, n.stream));
        HIPCHK(hipMemcpyAsync(n.R + off, f->Rqu[k].data(), np * ld * sizeof(float), hipMemcpyHostToDevice, n.stream));
        HIPCHK(hipMemcpyAsync(n.DS + off, f->DSqu[k].data(), np * ld * sizeof(float), hipMemcpyHostToDevice,
                              n.stream));
      }
    HIPCHK(hipMemcpyAsync(n.fluxes, flux.data(), flux.size() * sizeof(float), hipMemcpyHostToDevice, n.stream));
    // this iteration's radiance truncation (RADIANCE_TRUNCATION follows
    // SPLIT_GRID, shdomsub1.f:723-731) with the FIXSH of the last one
    launch_radiance_truncation(n.stream, n.R, n.ld, n.tc, n.ns, n.lofj, n.nr_next, q.npts, q.ml, q.mm, q.nlm,
                               q.d.shacc, s.nr_next_fixsh ? 1 : 0, q.d.high_order_radiance ? 1 : 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(n.stream));
    n.initialized = s.initialized;
    n.iters = s.iters;
    n.iters_last = s.iters_last;
    n.nsweep_launch = s.nsweep_launch;
    n.solcrit = s.solcrit;
    n.deljdot = s.deljdot;
    n.deljold = s.deljold;
    n.deljnew = s.deljnew;
    n.jnorm = s.jnorm;
    n.nr_next_fixsh = s.nr_next_fixsh;
    n.splitcrit = s.splitcrit;
    for (int k = 0; k < 8; k++) n.tsum[k] = s.tsum[k];
  }
  n.own_plan = owned;
  state_free(s);
  s = n;
  return AT3D_OK;
}

int state_reset_grid(State& s, std::string& err) {
  if (!s.own_plan) return AT3D_OK;
  return state_replace(s, nullptr, nullptr, err);
}

int state_set_grid(State& s, int npts, int ncells, const float* gridpos, const int32_t* gridptr,
                   const int32_t* neighptr, const int32_t* treeptr, const int8_t* cellflags, std::string& err) {
  const Plan& b = *s.base_plan;
  bool base = npts == b.npts && ncells == b.ncells;
  for (size_t i = 0; base && i < (size_t)ncells * 2; i++) base = treeptr[i] == -1;
  if (base) {
    // the base grid itself (a saved solution of an unrefined solve)
    base = memcmp(gridpos, b.gridpos.data(), (size_t)npts * 3 * sizeof(float)) == 0 &&
           memcmp(gridptr, b.gridptr.data(), (size_t)ncells * 8 * sizeof(int32_t)) == 0 &&
           memcmp(neighptr, b.neighptr.data(), (size_t)ncells * 6 * sizeof(int32_t)) == 0;
    if (!base) {
      err = "loaded grid: base grid differs from this solver's";
      return AT3D_ERR_ARG;
    }
    return state_reset_grid(s, err);
  }
  auto q = std::make_shared<Plan>(b);
  q->medium_on_device = false;   // the loaded grid's medium is made on the host
  q->host_medium = true;
  TRY(plan_set_grid(*q, npts, ncells, gridpos, gridptr, neighptr, treeptr, cellflags, err));
  return state_replace(s, q, nullptr, err);
}

static SourceArgs source_args(State& s, int first, int fixsh) {
  const Plan& p = *s.plan;
  SourceArgs a;
  a.R = s.R;
  a.nr = s.nr;
  a.nr_next = s.nr_next;
  a.tc = s.tc;
  a.radmin = p.d.shacc;
  a.highorder = p.d.high_order_radiance ? 1 : 0;
  a.alb = s.alb;
  a.chi = s.chi;
  a.f0 = s.f0;
  a.ysun = s.ysun;
  a.lofj = s.lofj;
  a.S = s.S;
  a.DS = s.DS;
  a.ns = s.ns;
  a.nos = s.nos;
  a.partial = s.partial;
  a.ld = s.ld;
  a.npts = p.npts;
  a.ml = p.ml;
  a.mm = p.mm;
  a.nlm = p.nlm;
  a.shacc = p.d.shacc;
  a.first = first;
  a.accel = p.d.accelflag ? 1 : 0;
  a.fixsh = fixsh;
  a.nst = s.nst;
  a.psh = s.psh;
  a.chip = s.chip;
  return a;
}

// INIT_SOLUTION's radiance and source (shdomsub1.f:113-442): the two-stream
// radiance (four terms, NR = 4) and the first COMPUTE_SOURCE.  NS and NOS keep
// the extents of the terms S and DS hold from an earlier solve, which the
// first source pass clears.
int state_init_solution(State& s, std::string& err) {
  // a state on a refined grid starts again from its base grid
  if (s.own_plan) TRY(state_replace(s, nullptr, nullptr, err));
  const Plan& p = *s.plan;
  launch_init_radiance(s.stream, s.R, s.R0, s.nr, s.ld, p.npts);
  // the two-stream start is unpolarized: Q and U start at zero
  if (s.nst == 3) HIPCHK(hipMemsetAsync(s.R + s.psh, 0, 2 * s.psh * sizeof(float), s.stream));
  launch_compute_source(s.stream, source_args(s, 1, 0), s.nsrc_blocks);
  s.nr_next_fixsh = false;
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s.stream));
  s.solcrit = 1.0f;
  s.iters = 0;
  s.initialized = true;
  return AT3D_OK;
}

// Same base grid (sizes, boundaries and the coordinates bit for bit) and the
// same spherical-harmonic layout (ML, MM), so that point i and term J mean
// the same in both states.
static bool same_grid_and_layout(const Plan& p, const Plan& q) {
  return p.npts == q.npts && p.ncells == q.ncells && p.nx1 == q.nx1 && p.ny1 == q.ny1 &&
         p.nz == q.nz && p.bcflag == q.bcflag && p.ipflag == q.ipflag && p.ml == q.ml &&
         p.mm == q.mm && p.nlm == q.nlm && p.xgrid == q.xgrid && p.ygrid == q.ygrid &&
         p.zgrid == q.zgrid && p.treeptr == q.treeptr &&
         (p.treeptr.empty() || (p.gridpos == q.gridpos && p.gridptr == q.gridptr));
}

// Warm start from another state's solution: its spherical-harmonic radiance
// (with its term counts) and hemispherical fluxes, as a saved solution
// carries them (a render before this state's first iteration takes the
// surface's downwelling flux from them), then the first source function of
// this state's medium.
int state_init_solution_from(State& s, State& src, std::string& err) {
  const Plan& p = *s.plan;
  const Plan& q = *src.plan;
  if (!src.initialized || !same_grid_and_layout(p, q) || s.ld != src.ld) {
    err = "warm start needs an initialised source state on the same grid and angular resolution";
    return AT3D_ERR_ARG;
  }
  if (s.nst != src.nst) {
    err = "warm start needs a source state with the same number of Stokes components";
    return AT3D_ERR_ARG;
  }
  const size_t nsh = (size_t)s.nptsp * s.ld * s.nst;
  HIPCHK(hipStreamSynchronize(src.stream));
  HIPCHK(hipMemcpyAsync(s.R, src.R, nsh * sizeof(float), hipMemcpyDeviceToDevice, s.stream));
  HIPCHK(hipMemcpyAsync(s.nr, src.nr, p.npts * sizeof(int), hipMemcpyDeviceToDevice, s.stream));
  HIPCHK(hipMemcpyAsync(s.fluxes, src.fluxes, 2 * (size_t)p.npts * sizeof(float),
                        hipMemcpyDeviceToDevice, s.stream));
  launch_compute_source(s.stream, source_args(s, 1, 0), s.nsrc_blocks);
  s.nr_next_fixsh = false;
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s.stream));
  s.solcrit = 1.0f;
  s.iters = 0;
  s.initialized = true;
  return AT3D_OK;
}

// The path-integration sweeps of one iteration: downwelling levels, surface
// boundary, upwelling levels.
static void enqueue_sweeps(State& s, const int* stop = nullptr) {
  const Plan& p = *s.plan;
  SweepArgs a = sweep_args(s);
  a.stop = stop;
  const float albpi = p.d.gndalbedo / acosf(-1.0f);
  const bool ring = ring_active(s);
  // the ring sweep writes the boundary radiances itself; events around its
  // launches time the sweep kernel alone
  // which ring instantiations each hemisphere's columns need
  auto needs = [&](int c0, int c1, bool& fast, bool& full) {
    fast = full = false;
    for (int c = c0; c < c1; c++) {
      const bool f = s.ring_full || s.col_general_h[c] != 0;
      full = full || f;
      fast = fast || !f;
    }
  };
  bool fast = false, full = false;
  // With NSTOKES 3 only the I sweep of the upwelling hemisphere waits for the
  // surface (Q and U see a zero boundary), so the first launch also runs the
  // upwelling Q and U columns: 5 of the 6 (hemisphere, component) sweeps in
  // one grid, which fills the CUs better than two grids of three
  const bool merge_qu = ring && s.nst == 3;
  if (ring) {
    // the first launch is timed from the event the caller records just
    // before this function, the second to the one just after it
    if (merge_qu) {
      SweepArgs a1 = a;
      a1.qu_only_from = s.ndown;
      needs(0, p.nang, fast, full);
      launch_sweep_ring(s.stream, a1, 0, p.nang, s.ring_lds, s.ring_lds_full, fast, full);
    } else {
      needs(0, s.ndown, fast, full);
      launch_sweep_ring(s.stream, a, 0, s.ndown, s.ring_lds, s.ring_lds_full, fast, full);
    }
    (void)hipEventRecord(s.ev_sweep[1], s.stream);
  } else {
    launch_sweep_preset(s.stream, a, false, 0, s.ndown);
    launch_sweep_columns(s.stream, a, 0, s.ndown);
  }
  BottomRecs br{ring ? s.rec : nullptr, p.npts, s.ndown, p.nang, s.col_oct, s.oct_pos};
  launch_bottom_bc(s.stream, s.dorad, s.nptsp, s.bcbot, (int)p.bcbot.size(), s.col_of_ref_d, s.fluxw,
                   s.ndown, s.dirflux, albpi, s.bcrad_bot, br, stop);
  if (ring) {
    needs(s.ndown, p.nang, fast, full);
    (void)hipEventRecord(s.ev_sweep[2], s.stream);
    SweepArgs a2 = a;
    if (merge_qu) a2.nst = 1;   // the upwelling I columns alone
    launch_sweep_ring(s.stream, a2, s.ndown, p.nang - s.ndown, s.ring_lds, s.ring_lds_full, fast, full);
  } else {
    launch_sweep_preset(s.stream, a, true, s.ndown, p.nang - s.ndown);
    launch_sweep_columns(s.stream, a, s.ndown, p.nang - s.ndown);
  }
}

// The point-ordered sweep's level lists (every octant's non-preset points
// by level, their start cells and the level offsets), uploaded on its first use.
static int ensure_point_order(State& s, std::string& err) {
  if (s.lvl_pts) return AT3D_OK;
  const Plan& p = *s.plan;
  std::vector<int32_t> lp, lc, lo;
  for (int o = 0; o < 8; o++) {
    const OctantPlan& op = p.oct[o];
    const int base = (int)lp.size();
    lp.insert(lp.end(), op.lvl_pts.begin(), op.lvl_pts.end());
    lc.insert(lc.end(), op.lvl_cell.begin(), op.lvl_cell.end());
    for (int L = 0; L <= op.nlevels; L++) lo.push_back(base + op.lvl_off[L]);
  }
  TRY(upload(s, s.lvl_pts, lp.data(), lp.size(), err));
  TRY(upload(s, s.lvl_cell, lc.data(), lc.size(), err));
  TRY(upload(s, s.lvl_off, lo.data(), lo.size(), err));
  return AT3D_OK;
}

static int run_sweeps(State& s, std::string& err) {
  if (!ring_active(s)) TRY(ensure_point_order(s, err));
  // the level-ordered sweep is three launches per pass (and is timed by
  // events between them); the point-ordered one is replayed from a graph
  if (!s.use_graph || ring_active(s)) {
    enqueue_sweeps(s);
    HIPCHK(hipGetLastError());
    return AT3D_OK;
  }
  const int ring = ring_active(s) ? 1 : 0;
  if (!s.sweep_exec || s.graph_general != s.force_general || s.graph_ring != ring) {
    if (s.sweep_exec) (void)hipGraphExecDestroy(s.sweep_exec);
    if (s.sweep_graph) (void)hipGraphDestroy(s.sweep_graph);
    s.sweep_exec = nullptr;
    s.sweep_graph = nullptr;
    std::lock_guard<std::mutex> lock(device_mutex());
    HIPCHK(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
    enqueue_sweeps(s);
    hipError_t e1 = hipGetLastError();
    hipError_t e2 = hipStreamEndCapture(s.stream, &s.sweep_graph);
    if (e1 != hipSuccess || e2 != hipSuccess) {
      err = std::string("sweep graph capture failed: ") +
            hipGetErrorString(e1 != hipSuccess ? e1 : e2);
      return AT3D_ERR_DEVICE;
    }
    HIPCHK(hipGraphInstantiate(&s.sweep_exec, s.sweep_graph, nullptr, nullptr, 0));
    s.graph_general = s.force_general;
    s.graph_ring = ring;
  }
  HIPCHK(hipGraphLaunch(s.sweep_exec, s.stream));
  return AT3D_OK;
}

// CALC_ACCEL_SOLCRIT (shdom_nompi.f:317-349) in single precision.
static void accel_solcrit(bool doaccel, float deljdot, float deljold, float deljnew, float jnorm,
                          float& a, float& accelpar, float& solcrit) {
  if (doaccel && a == 0.0f && deljnew < deljold) {
    float r = sqrtf(deljnew / deljold);
    float c = deljdot / sqrtf(deljold * deljnew);
    c = std::max(-1.0f, std::min(1.0f, c));
    float theta = acosf(c);
    a = (1 - r * cosf(theta) + powf(r, 1 + 0.5f * 3.14159f / theta)) /
            (1 + r * r - 2 * r * cosf(theta)) - 1.0f;
    a = std::min(10.0f, std::max(0.0f, a));
  } else {
    a = 0.0f;
  }
  accelpar = a;
  if (jnorm > 0.0f)
    solcrit = sqrtf(deljnew / jnorm);
  else if (deljnew == 0.0f)
    solcrit = 0.0f;
}

// The Q and U rows of SH->DO (in the layout the I rows have: each octant's
// level order for the ring sweep, point order otherwise) and the Q and U
// planes of DO->SH of a polarized state; nothing for NSTOKES = 1.
static void enqueue_sh_to_do_qu(State& s, const int* stop) {
  if (s.nst != 3) return;
  const Plan& p = *s.plan;
  if (ring_active(s))
    launch_sh_to_do_perm_qu(s.stream, s.S + s.psh, s.psh, s.Yqu, s.ext, s.ns, s.oct_perm, s.sh_tiles_qu,
                            2 * s.n_sh_tiles, s.dosrc + s.pdo, s.nptsp, s.ld, s.nptsp, p.npts, stop);
  else
    launch_sh_to_do_qu(s.stream, s.S + s.psh, s.psh, s.Yqu, s.ext, s.dosrc + s.pdo, s.nptsp, 2 * s.ldo, s.ld,
                       s.nptsp, p.npts, stop);
}

static void enqueue_do_to_sh_qu(State& s, const int* stop) {
  if (s.nst != 3) return;
  const Plan& p = *s.plan;
  launch_do_to_sh_qu(s.stream, s.dorad + s.pdo, s.Wqu, s.nr, s.R + s.psh, s.psh, s.nptsp, s.ld, s.ldo, s.nptsp,
                     p.npts, stop);
}

// One iteration under device control (IterCtl): the loop test and FIXSH,
// SH->DO, the sweeps, DO->SH, the source pass with the next radiance
// truncation, the reduction, SOLCRIT and the acceleration; every kernel
// returns at once once the control block says stop.  The NR buffers are
// those the host pointers name now.  (Inside a stream capture an event
// record only marks a dependency: kernel times come from the iterations
// launched one by one.)
static void enqueue_iteration(State& s) {
  const Plan& p = *s.plan;
  const int* stop = &s.ctl->stop;
  launch_iter_begin(s.stream, s.ctl);
  if (ring_active(s))
    launch_sh_to_do_perm(s.stream, s.S, s.Y, s.ext, s.ns, s.oct_perm, s.sh_tiles, s.n_sh_tiles, s.dosrc,
                         s.nptsp, s.ld, s.nptsp, p.npts, stop);
  else
    launch_sh_to_do(s.stream, s.S, s.Y, s.ext, s.dosrc, s.nptsp, s.ldo, s.ld, s.nptsp, p.npts, p.nang, stop);
  enqueue_sh_to_do_qu(s, stop);
  enqueue_sweeps(s, stop);
  launch_do_to_sh(s.stream, s.dorad, s.W, s.nr, s.R, s.nptsp, s.ld, s.ldo, s.nptsp, p.npts, stop);
  enqueue_do_to_sh_qu(s, stop);
  SourceArgs sa = source_args(s, 0, 0);
  sa.ctl = s.ctl;
  launch_compute_source(s.stream, sa, s.nsrc_blocks);
  launch_reduce_dots(s.stream, s.partial, s.nsrc_blocks, s.dots, stop);
  launch_solcrit(s.stream, s.ctl, s.dots, s.errflag);
  launch_accelerate(s.stream, s.S, s.DS, s.ns, s.nos, s.ld, p.npts, 0.0f, s.ctl, s.nst, s.psh);
}

// The graph of ITER_BATCH iterations for the NR buffer parity k (0: s.nr is
// the first buffer), captured on the state's stream.
static int capture_iterations(State& s, int k, std::string& err) {
  if (s.iter_exec[k]) (void)hipGraphExecDestroy(s.iter_exec[k]);
  if (s.iter_graph[k]) (void)hipGraphDestroy(s.iter_graph[k]);
  s.iter_exec[k] = nullptr;
  s.iter_graph[k] = nullptr;
  std::lock_guard<std::mutex> lock(device_mutex());
  HIPCHK(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
  for (int b = 0; b < State::ITER_BATCH; b++) {
    std::swap(s.nr, s.nr_next);
    enqueue_iteration(s);
  }
  hipError_t e1 = hipGetLastError();
  hipError_t e2 = hipStreamEndCapture(s.stream, &s.iter_graph[k]);
  // ITER_BATCH is even: the host pointers are back where they were
  if (e1 != hipSuccess || e2 != hipSuccess) {
    err = std::string("iteration graph capture failed: ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2);
    return AT3D_ERR_DEVICE;
  }
  HIPCHK(hipGraphInstantiate(&s.iter_exec[k], s.iter_graph[k], nullptr, nullptr, 0));
  s.iter_key[k] = (ring_active(s) ? 1 : 0) | (s.force_general ? 2 : 0) | (s.ring_full ? 4 : 0);
  return AT3D_OK;
}

// SOLUTION_ITERATIONS from captured graphs: the host launches a batch of
// iterations and reads the control block once per batch; the device decides
// when the solve stops.
static int state_solve_graph(State& s, int maxiter, std::string& err) {
  const Plan& p = *s.plan;
  IterCtl& c = *s.ctl_h;
  memset(&c, 0, sizeof c);
  c.total = s.iters;
  c.solcrit = s.solcrit;
  c.maxiter = maxiter;
  c.iterfixsh = p.d.iterfixsh;
  c.accel = p.d.accelflag ? 1 : 0;
  c.one_iter = s.albmax < p.d.solacc ? 1 : 0;
  c.solacc = p.d.solacc;
  HIPCHK(hipMemcpyAsync(s.ctl, &c, sizeof c, hipMemcpyHostToDevice, s.stream));
  const int parity = s.nr == s.nr_buf0 ? 0 : 1;
  s.dosrc_level_order = ring_active(s);
  const int key = (ring_active(s) ? 1 : 0) | (s.force_general ? 2 : 0) | (s.ring_full ? 4 : 0);
  if (!s.iter_exec[parity] || s.iter_key[parity] != key) TRY(capture_iterations(s, parity, err));
  // every batch runs ITER_BATCH iterations unless the device stops: the
  // host bound only guards against a control block that never moves
  const int max_batches = (maxiter - s.iters) / State::ITER_BATCH + 2;
  for (int batch = 0;; batch++) {
    if (batch >= max_batches) {
      char buf[200];
      snprintf(buf, sizeof buf,
               "iteration graph: the device control block did not reach its stop condition "
               "(stop %d, iterations %d, total %d, SOLCRIT %g after %d batches)",
               c.stop, c.iters, c.total, (double)c.solcrit, batch);
      err = buf;
      return AT3D_ERR_DEVICE;
    }
    HIPCHK(hipGraphLaunch(s.iter_exec[parity], s.stream));
    HIPCHK(hipMemcpyAsync(&c, s.ctl, sizeof c, hipMemcpyDeviceToHost, s.stream));
    HIPCHK(hipStreamSynchronize(s.stream));
    s.iters_last = c.iters;
    if (c.stop || c.total >= maxiter || c.solcrit <= p.d.solacc) break;
  }
  if (c.iters & 1) std::swap(s.nr, s.nr_next);
  s.iters = c.total;
  s.solcrit = c.solcrit;
  s.deljdot = c.deljdot;
  s.deljold = c.deljold;
  s.deljnew = c.deljnew;
  s.jnorm = c.jnorm;
  s.nr_next_fixsh = c.fixsh != 0;
  s.nsweep_launch = c.iters;
  if (c.stop == 2) {
    int ef = 0;
    HIPCHK(hipMemcpy(&ef, s.errflag, sizeof(int), hipMemcpyDeviceToHost));
    char buf[160];
    snprintf(buf, sizeof buf, "BACK_INT_GRID: invalid characteristic (flags 0x%x) at iteration %d", ef,
             s.iters);
    err = buf;
    return AT3D_ERR_NUMERIC;
  }
  return AT3D_OK;
}

// One iteration under host control (the loop body of SOLUTION_ITERATIONS,
// shdomsub1.f:723-790, after the splitting): the NR made by the last source
// pass, SH->DO, the sweeps, DO->SH, the source with the next radiance
// truncation, and the dots read back for SOLCRIT and the acceleration.
static int host_iteration(State& s, bool& fixsh, float& a, int verbose, std::string& err) {
  const Plan& p = *s.plan;
  s.iters++;
  s.iters_last++;
  // RADIANCE_TRUNCATION of this iteration was done by the last source pass
  // (into nr_next): it becomes the current NR
  std::swap(s.nr, s.nr_next);
  (void)hipEventRecord(s.ev[0], s.stream);
  s.dosrc_level_order = ring_active(s);
  if (s.dosrc_level_order)
    launch_sh_to_do_perm(s.stream, s.S, s.Y, s.ext, s.ns, s.oct_perm, s.sh_tiles, s.n_sh_tiles,
                         s.dosrc, s.nptsp, s.ld, s.nptsp, p.npts);
  else
    launch_sh_to_do(s.stream, s.S, s.Y, s.ext, s.dosrc, s.nptsp, s.ldo, s.ld, s.nptsp, p.npts,
                    p.nang);
  enqueue_sh_to_do_qu(s, nullptr);
  (void)hipEventRecord(s.ev[2], s.stream);
  TRY(run_sweeps(s, err));
  (void)hipEventRecord(s.ev[3], s.stream);
  launch_do_to_sh(s.stream, s.dorad, s.W, s.nr, s.R, s.nptsp, s.ld, s.ldo, s.nptsp, p.npts);
  enqueue_do_to_sh_qu(s, nullptr);
  (void)hipEventRecord(s.ev[4], s.stream);
  if (s.solcrit < 0.001f || s.iters > p.d.iterfixsh) fixsh = true;
  launch_compute_source(s.stream, source_args(s, 0, fixsh ? 1 : 0), s.nsrc_blocks);
  s.nr_next_fixsh = fixsh;
  launch_reduce_dots(s.stream, s.partial, s.nsrc_blocks, s.dots);
  (void)hipEventRecord(s.ev[5], s.stream);
  HIPCHK(hipGetLastError());
  // the four dot products and the error flag into pinned host memory
  double* dots = s.host_dots;
  int* efp = reinterpret_cast<int*>(s.host_dots + 4);
  HIPCHK(hipMemcpyAsync(dots, s.dots, 4 * sizeof(double), hipMemcpyDeviceToHost, s.stream));
  HIPCHK(hipMemcpyAsync(efp, s.errflag, sizeof(int), hipMemcpyDeviceToHost, s.stream));
  HIPCHK(hipStreamSynchronize(s.stream));
  const int ef = *efp;
  if (ef) {
    char buf[160];
    snprintf(buf, sizeof buf, "BACK_INT_GRID: invalid characteristic (flags 0x%x) at iteration %d",
             ef, s.iters);
    err = buf;
    return AT3D_ERR_NUMERIC;
  }
  // (the radiance truncation is part of the source pass: tsum[0] stays 0)
  float ms[5] = {};
  for (int e = 1; e < 5; e++) (void)hipEventElapsedTime(&ms[e], s.ev[e == 1 ? 0 : e], s.ev[e + 1]);
  s.tsum[1] += ms[1];  // SH -> DO
  s.tsum[2] += ms[2];  // sweeps (both hemispheres, surface boundary)
  s.tsum[3] += ms[3];  // DO -> SH
  s.tsum[4] += ms[4];  // source (with the next truncation) + reduction
  if (ring_active(s)) {
    float a0 = 0.f, a1 = 0.f;
    (void)hipEventElapsedTime(&a0, s.ev[2], s.ev_sweep[1]);
    (void)hipEventElapsedTime(&a1, s.ev_sweep[2], s.ev[3]);
    s.tsum[6] += a0 + a1;  // the two sweep kernel launches alone
  }
  s.nsweep_launch += 1;
  s.deljdot = (float)dots[0];
  s.deljold = (float)dots[1];
  s.deljnew = (float)dots[2];
  s.jnorm = (float)dots[3];
  float accelpar;
  accel_solcrit(p.d.accelflag != 0, s.deljdot, s.deljold, s.deljnew, s.jnorm, a, accelpar, s.solcrit);
  if (accelpar > 0.0f)
    launch_accelerate(s.stream, s.S, s.DS, s.ns, s.nos, s.ld, p.npts, accelpar, nullptr, s.nst, s.psh);
  if (s.albmax < p.d.solacc) s.solcrit = p.d.solacc;
  if (verbose) {
    printf("  %4d %8.3f %10.3e %8d\n", s.iters, log10f(std::max(s.solcrit, 1.0e-20f)), s.splitcrit, p.npts);
    fflush(stdout);
  }
  return AT3D_OK;
}

// The splitting test of one iteration (SPLIT_GRID, shdomsub1.f:694-715): the
// source (and, to divide cells, the radiance) comes to the host, the grid is
// refined there, and the state moves to the refined grid when it changed.
static int split_step(State& s, bool dosplit, SplitSchedule& sc, std::string& err) {
  const Plan& p = *s.plan;
  TRY(ensure_host_medium(p, err));   // the splitting reads and extends the host optics
  const size_t np = p.npts, ld = s.ld;
  SplitFields f;
  f.ld = s.ld;
  f.nst = s.nst;
  f.S.resize(np * ld);
  f.ns.resize(np);
  HIPCHK(hipMemcpyAsync(f.S.data(), s.S, np * ld * sizeof(float), hipMemcpyDeviceToHost, s.stream));
  HIPCHK(hipMemcpyAsync(f.ns.data(), s.ns, np * sizeof(int), hipMemcpyDeviceToHost, s.stream));
  if (dosplit) {
    f.R.resize(np * ld);
    f.DS.resize(np * ld);
    f.nr.resize(np);
    f.nos.resize(np);
    HIPCHK(hipMemcpyAsync(f.R.data(), s.R, np * ld * sizeof(float), hipMemcpyDeviceToHost, s.stream));
    HIPCHK(hipMemcpyAsync(f.DS.data(), s.DS, np * ld * sizeof(float), hipMemcpyDeviceToHost, s.stream));
    HIPCHK(hipMemcpyAsync(f.nr.data(), s.nr, np * sizeof(int), hipMemcpyDeviceToHost, s.stream));
    HIPCHK(hipMemcpyAsync(f.nos.data(), s.nos, np * sizeof(int), hipMemcpyDeviceToHost, s.stream));
    if (s.nst == 3)
      for (int k = 0; k < 2; k++) {
        const size_t off = (size_t)(k + 1) * s.psh;   // the Q and U planes
        f.Squ[k].resize(np * ld);
        f.Rqu[k].resize(np * ld);
        f.DSqu[k].resize(np * ld);
        HIPCHK(hipMemcpyAsync(f.Squ[k].data(), s.S + off, np * ld * sizeof(float), hipMemcpyDeviceToHost, s.stream));
        HIPCHK(hipMemcpyAsync(f.Rqu[k].data(), s.R + off, np * ld * sizeof(float), hipMemcpyDeviceToHost, s.stream));
        HIPCHK(hipMemcpyAsync(f.DSqu[k].data(), s.DS + off, np * ld * sizeof(float), hipMemcpyDeviceToHost,
                              s.stream));
      }
  }
  HIPCHK(hipStreamSynchronize(s.stream));
  if (!dosplit) {
    sc.splitcrit = split_criterion(p, f);
    s.splitcrit = sc.splitcrit;
    return AT3D_OK;
  }
  auto q = std::make_shared<Plan>(p);
  q->medium_on_device = false;   // the refined grid's medium is made on the host
  bool outofmem = sc.outofmem;
  float crit = 0.0f;
  TRY(split_grid(*q, f, true, outofmem, sc.cursplitacc, crit, err));
  sc.outofmem = outofmem;
  sc.splitcrit = crit;
  s.splitcrit = crit;
  if (q->ncells == p.ncells) return AT3D_OK;   // nothing was divided
  TRY(plan_refresh_points(*q, err));
  return state_replace(s, q, &f, err);
}

// SOLUTION_ITERATIONS with adaptive cell splitting (shdomsub1.f:636-790):
// the splitting accuracy starts at SPLITACC * ADAPTRANGE and follows the
// geometric mean of SOLCRIT down to SPLITACC while SOLCRIT falls from 0.1 to
// 0.001; the loop also runs on while cells remain above the splitting
// accuracy.  Iterations run under host control (the grid can change between
// them), the level-ordered sweep stays off (it needs the base grid).
static int solve_split(State& s, int maxiter, int verbose, std::string& err) {
  const float splitacc = s.plan->d.splitacc, solacc = s.plan->d.solacc;
  const float endadaptsol = 0.001f, startadaptsol = 0.1f;
  const float adaptrange = startadaptsol / (3.0f * endadaptsol);
  SplitSchedule sc;
  sc.cursplitacc = splitacc * adaptrange;
  sc.startsplitacc = sc.cursplitacc;
  sc.avgsolcrit = s.solcrit;
  sc.splitcrit = 0.0f;
  float a = 0.0f;
  bool fixsh = false;
  while (s.iters < maxiter &&
         (s.solcrit > solacc || (sc.splitcrit > splitacc && sc.cursplitacc > splitacc && !sc.outofmem))) {
    sc.avgsolcrit = sqrtf(sc.avgsolcrit * s.solcrit);
    const bool dosplit =
        s.solcrit <= startadaptsol && (s.solcrit > endadaptsol || sc.cursplitacc > splitacc) && !sc.outofmem;
    const float beta = logf(sc.startsplitacc / splitacc) / logf(adaptrange);
    sc.cursplitacc = std::min(sc.cursplitacc,
                              std::max(splitacc, splitacc * powf(sc.avgsolcrit / (3.0f * endadaptsol), beta)));
    if (s.solcrit <= endadaptsol) sc.cursplitacc = splitacc;
    if (sc.testing) {
      TRY(split_step(s, dosplit, sc, err));
      if (s.solcrit > startadaptsol) sc.startsplitacc = sc.splitcrit;
    }
    if (s.solcrit <= endadaptsol) sc.testing = false;
    TRY(host_iteration(s, fixsh, a, verbose, err));
  }
  const Plan& p = *s.plan;
  if (s.iters_last > 0)
    launch_fluxes(s.stream, s.dorad, s.nptsp, s.col_of_ref_d, s.fluxw, s.ndown, p.nang, p.npts, s.fluxes);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s.stream));
  return AT3D_OK;
}

int state_solve(State& s, int maxiter, int verbose, std::string& err) {
  if (!s.initialized) TRY(state_init_solution(s, err));
  const Plan& p = *s.plan;
  HIPCHK(hipMemsetAsync(s.errflag, 0, sizeof(int), s.stream));
  float a = 0.0f;
  bool fixsh = false;
  for (auto& t : s.tsum) t = 0.0;
  s.last_solve_graph = false;
  s.iters_last = 0;
  s.nsweep_launch = 0;
  // the NR waiting in nr_next was made with the last call's FIXSH; a call
  // starts with FIXSH off, so it is remade when that differs
  if (s.nr_next_fixsh && s.iters < maxiter && s.solcrit > p.d.solacc) {
    launch_radiance_truncation(s.stream, s.R, s.ld, s.tc, s.ns, s.lofj, s.nr_next, p.npts, p.ml, p.mm,
                               p.nlm, p.d.shacc, 0, p.d.high_order_radiance ? 1 : 0);
    s.nr_next_fixsh = false;
  }
  // (p is not used past this call: the state's plan changes when it splits)
  if (p.d.splitacc > 0.0f) return solve_split(s, maxiter, verbose, err);
  // graphs replay the level-ordered sweep; the point-ordered one (grids whose
  // ring does not fit in LDS, and the test options) iterates from the host
  if (s.use_iter_graph && !verbose && ring_active(s) && s.iters < maxiter && s.solcrit > p.d.solacc) {
    TRY(state_solve_graph(s, maxiter, err));
    s.last_solve_graph = true;
    launch_fluxes(s.stream, s.dorad, s.nptsp, s.col_of_ref_d, s.fluxw, s.ndown, p.nang, p.npts, s.fluxes);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s.stream));
    return AT3D_OK;
  }
  while (s.iters < maxiter && s.solcrit > p.d.solacc) TRY(host_iteration(s, fixsh, a, verbose, err));
  // hemispherical fluxes of the last path integration (a call that ran no
  // iteration keeps the ones the state holds)
  if (s.iters_last > 0)
    launch_fluxes(s.stream, s.dorad, s.nptsp, s.col_of_ref_d, s.fluxw, s.ndown, p.nang, p.npts, s.fluxes);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s.stream));
  return AT3D_OK;
}

}  // namespace at3d
