// Discrete-ordinate sweeps (BACK_INT_GRID3D_UNPOL, shdomsub1.f:3695-4032).
//
// The reference integrates one ordinate at a time, visiting grid points in its
// sweep order (SWEEPING_ORDER :3261) and tracing each point back to the first
// cell face whose four corner radiances are already known.  The host planner
// replays that order and assigns every point a dependency level per octant
// (all upwind face points of every ordinate of the octant sit on lower
// levels).  On the GPU one workgroup integrates one ordinate over the whole
// grid, walking the levels of the ordinate's octant with a workgroup barrier
// between levels, so a hemisphere's sweep is one launch whose workgroups run
// independently.
//
// k_sweep_ring is the production sweep.  Everything about a one-cell
// characteristic that stays fixed during a solve (face points, bilinear
// weights, transmission, the source-integral coefficients) is computed once
// per solve by k_sweep_records and stored in level order, so a level is read
// as one contiguous block.  The radiances and source values of the last
// RING_LEVELS levels and of the boundary points live in LDS; a level then
// costs LDS gathers, a few FMAs and a barrier that waits for LDS traffic only,
// while the next level's records stream in behind it.  Multi-cell traces
// (periodic seams, faces older than the ring) take the general trace through
// global memory behind a full barrier.  k_sweep_columns is the original
// sweep on point-ordered DO arrays; it serves grids whose ring does not fit
// in LDS and the general-trace check.
#include <hip/hip_runtime.h>
#include <limits.h>

#include "kernels.h"
#include "ring_plan.h"

namespace at3d {

// The part of a one-cell characteristic that does not depend on radiances
// computed during the sweep: the upwind face points and bilinear weights, and
// the cell integral of the source (radv = src*abs, transmission).
struct SimpleStep {
  int i[4];
  double f[4];
  double radv, transmit;
};

__device__ inline void sweep_simple_static(const SweepArgs& a, const OrdGeom& og, int col, int pt,
                                           SimpleStep& st, int kst = 0) {
#pragma clang fp contract(off)
  const float* xg = a.xgrid;
  const float* yg = a.ygrid;
  const float* zg = a.zgrid;
  const int nyz = a.nz * a.ny1;
  const int ix = pt / nyz, rem = pt - ix * nyz;
  const int iy = rem / a.nz, iz = rem - iy * a.nz;
  const int ixf = og.bitx ? ix - 1 : ix + 1;
  const int iyf = og.bity ? iy - 1 : iy + 1;
  const int izf = og.bitz ? iz - 1 : iz + 1;
  double xe = xg[ix], ye = yg[iy], ze = zg[iz];
  const double sox = ((double)xg[ixf] - xe) * og.cxinv;
  const double soy = ((double)yg[iyf] - ye) * og.cyinv;
  const double soz = ((double)zg[izf] - ze) * og.czinv;
  double so = sox < soy ? sox : soy;
  so = so < soz ? so : soz;
  xe = xe + so * og.cx;
  ye = ye + so * og.cy;
  ze = ze + so * og.cz;
  const int x0 = min(ix, ixf), x1 = max(ix, ixf);
  const int y0 = min(iy, iyf), y1 = max(iy, iyf);
  const int z0 = min(iz, izf), z1 = max(iz, izf);
  const int nz = a.nz, ny1 = a.ny1;
  auto P = [nz, ny1](int x, int y, int z) { return z + nz * (y + ny1 * x); };
  int i1, i2, i3, i4;
  double u, v;
  if (sox <= soz && sox <= soy) {
    i1 = P(ixf, y0, z0); i2 = P(ixf, y1, z0); i3 = P(ixf, y0, z1); i4 = P(ixf, y1, z1);
    u = (ze - (double)zg[z0]) / ((double)zg[z1] - (double)zg[z0]);
    v = (ye - (double)yg[y0]) / ((double)yg[y1] - (double)yg[y0]);
  } else if (soy <= soz) {
    i1 = P(x0, iyf, z0); i2 = P(x1, iyf, z0); i3 = P(x0, iyf, z1); i4 = P(x1, iyf, z1);
    u = (ze - (double)zg[z0]) / ((double)zg[z1] - (double)zg[z0]);
    v = (xe - (double)xg[x0]) / ((double)xg[x1] - (double)xg[x0]);
  } else {
    i1 = P(x0, y0, izf); i2 = P(x1, y0, izf); i3 = P(x0, y1, izf); i4 = P(x1, y1, izf);
    u = (ye - (double)yg[y0]) / ((double)yg[y1] - (double)yg[y0]);
    v = (xe - (double)xg[x0]) / ((double)xg[x1] - (double)xg[x0]);
  }
  const float* ext = a.ext;
  const float* se = a.srcext + kst * a.pdo + (size_t)col * a.ldp;
  const double f1 = (1 - u) * (1 - v), f2 = (1 - u) * v, f3 = u * (1 - v), f4 = u * v;
  const double ext0 = f1 * ext[i1] + f2 * ext[i2] + f3 * ext[i3] + f4 * ext[i4];
  const double srcext0 = f1 * se[i1] + f2 * se[i2] + f3 * se[i3] + f4 * se[i4];
  double src, absc, trc;
  cell_integral(ext0, (double)ext[pt], srcext0, (double)se[pt], so, src, absc, trc, kst == 0);
  st.i[0] = i1;
  st.i[1] = i2;
  st.i[2] = i3;
  st.i[3] = i4;
  st.f[0] = f1;
  st.f[1] = f2;
  st.f[2] = f3;
  st.f[3] = f4;
  st.radv = 0.0 + 1.0 * src * absc;
  st.transmit = 1.0 * trc;
}

// Adds the transmitted upwind radiance (interpolated on the face) to a static step.
__device__ inline float sweep_simple_finish(const SimpleStep& st, const float* rad) {
#pragma clang fp contract(off)
  const double rad0 = st.f[0] * rad[st.i[0]] + st.f[1] * rad[st.i[1]] + st.f[2] * rad[st.i[2]] +
                      st.f[3] * rad[st.i[3]];
  return (float)(st.radv + st.transmit * rad0);
}

__device__ inline float sweep_simple(const SweepArgs& a, const OrdGeom& og, int col, int pt, int kst = 0) {
  SimpleStep st;
  sweep_simple_static(a, og, col, pt, st, kst);
  return sweep_simple_finish(st, a.rad + kst * a.pdo + (size_t)col * a.ldp);
}

// General characteristic: trace through cells until a face with known radiances.
// The source row `se` is indexed by point, or through `pos` (point -> position)
// when it is stored in an octant's level order.
__device__ inline float sweep_general(const SweepArgs& a, const OrdGeom& og, int col, int o, int pt,
                                      int icell, int L, const float* se, const int32_t* pos, int kst = 0) {
  const int presetbit = og.bitz ? 2 : 1;
  const int32_t* rank = a.rank + (size_t)o * a.npts;
  const int32_t* level = a.level + (size_t)o * a.npts;
  const int rp = rank[pt];
  const float* ext = a.ext;
  auto SE = [se, pos](int q) { return se[pos ? pos[q] : q]; };
  const float* rad = a.rad + kst * a.pdo + (size_t)col * a.ldp;
  double xe = a.g.gridpos[pt * 3], ye = a.g.gridpos[pt * 3 + 1], ze = a.g.gridpos[pt * 3 + 2];
  double ext1 = ext[pt], srcext1 = SE(pt);
  double radv = 0.0, transmit = 1.0;
  for (int step = 0; step <= a.g.ncells; step++) {
    TraceStep t;
    if (!trace_step(a.g, og, icell, xe, ye, ze, t)) {
      atomicOr(a.errflag, 1);
      return 0.0f;
    }
    const double f1 = (1 - t.u) * (1 - t.v), f2 = (1 - t.u) * t.v;
    const double f3 = t.u * (1 - t.v), f4 = t.u * t.v;
    const double ext0 = f1 * ext[t.i[0]] + f2 * ext[t.i[1]] + f3 * ext[t.i[2]] + f4 * ext[t.i[3]];
    const double srcext0 = f1 * SE(t.i[0]) + f2 * SE(t.i[1]) + f3 * SE(t.i[2]) + f4 * SE(t.i[3]);
    double src, absc, trc;
    cell_integral(ext0, ext1, srcext0, srcext1, t.so, src, absc, trc, kst == 0);
    radv = radv + transmit * src * absc;
    transmit = transmit * trc;
    bool valid = true;
    for (int c = 0; c < 4; c++) {
      const int q = t.i[c];
      valid = valid && ((a.preset[q] & presetbit) || rank[q] < rp);
    }
    if (t.next < 0 || valid) {
      if (!valid) {
        atomicOr(a.errflag, 2);
        return 0.0f;
      }
      for (int c = 0; c < 4; c++) {
        const int q = t.i[c];
        if (!(a.preset[q] & presetbit) && level[q] >= L) atomicOr(a.errflag, 4);
      }
      const double rad0 = f1 * rad[t.i[0]] + f2 * rad[t.i[1]] + f3 * rad[t.i[2]] + f4 * rad[t.i[3]];
      radv = radv + transmit * rad0;
      return (float)radv;
    }
    ext1 = ext0;
    srcext1 = srcext0;
    icell = t.next;
  }
  atomicOr(a.errflag, 8);
  return 0.0f;
}

constexpr int SWEEP_THREADS = 512;
constexpr int SWEEP_K = 4;   // items per thread whose static part is prepared ahead

// One ordinate (workgroup) through every level of its octant.  The static
// part of each one-cell step of level L+1 (face points, weights, source
// integral) is prepared while level L is finishing; after the barrier only the
// upwind radiances are gathered.  Multi-cell traces and levels larger than
// SWEEP_K * SWEEP_THREADS points take the unprepared path.
__global__ __launch_bounds__(SWEEP_THREADS) void k_sweep_columns(SweepArgs a, int col0) {
  if (a.stop && *a.stop) return;
  const int col = col0 + blockIdx.x;
  const int kst = blockIdx.y;   // Stokes component (Q, U: no clamp of the cell source)
  const int o = a.col_oct[col];
  const OrdGeom og = a.ordg[col];
  const int nlev = a.nlev[o];
  const int32_t* off = a.lvl_off + a.lo_base[o];
  float* rad = a.rad + kst * a.pdo + (size_t)col * a.ldp;
  const float* se_row = a.srcext + kst * a.pdo + (size_t)col * a.ldp;
  const int tid = threadIdx.x;
  const bool general = a.force_general != 0;
  SimpleStep st[SWEEP_K];
  int code[SWEEP_K];
  auto prepare = [&](int L) {
    const int beg = off[L - 1], end = off[L];
#pragma unroll
    for (int k = 0; k < SWEEP_K; k++) {
      const int j = beg + tid + k * SWEEP_THREADS;
      code[k] = (j < end) ? a.lvl_pts[j] : INT_MIN;
      if (code[k] >= 0 && !general) sweep_simple_static(a, og, col, code[k], st[k], kst);
    }
  };
  if (nlev >= 1) prepare(1);
  for (int L = 1; L <= nlev; L++) {
    __syncthreads();
    const int beg = off[L - 1], end = off[L];
#pragma unroll
    for (int k = 0; k < SWEEP_K; k++) {
      const int c = code[k];
      if (c == INT_MIN) continue;
      if (c >= 0 && !general) {
        rad[c] = sweep_simple_finish(st[k], rad);
      } else {
        const int j = beg + tid + k * SWEEP_THREADS;
        const int pt = c >= 0 ? c : ~c;
        rad[pt] = sweep_general(a, og, col, o, pt, a.lvl_cell[j], L, se_row, nullptr, kst);
      }
    }
    for (int j = beg + tid + SWEEP_K * SWEEP_THREADS; j < end; j += SWEEP_THREADS) {
      const int c = a.lvl_pts[j];
      if (c >= 0 && !general) {
        rad[c] = sweep_simple(a, og, col, c, kst);
      } else {
        const int pt = c >= 0 ? c : ~c;
        rad[pt] = sweep_general(a, og, col, o, pt, a.lvl_cell[j], L, se_row, nullptr, kst);
      }
    }
    if (L < nlev) prepare(L + 1);
  }
}

// ---------------------------------------------------------------------------
// Level-ordered sweep with an LDS ring
// ---------------------------------------------------------------------------

// Step classes.  Every point of one level whose first cell step starts in a
// cell with the same IP flags steps, for a given ordinate, through a cell of
// the same shape (the grid is a tensor product), so the step's face
// interpolation factors (u, v) and path length are those of the class; the
// table holds them for every (column, class), traced once from a
// representative point of the class.
__device__ inline int step_class(const SweepArgs& a, int pt, int8_t cellflags) {
  const int iz = pt - a.nz * (int)(((unsigned long long)pt * a.nz_magic) >> 40);
  return iz * 4 + (cellflags & 3);
}

__global__ __launch_bounds__(256) void k_sweep_geo(SweepArgs a, int ncols) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncols * a.ncls) return;
  const int col = i / a.ncls, cls = i - col * a.ncls;
  const int o = a.col_oct[col];
  const int2 rep = a.geo_rep[(size_t)o * a.ncls + cls];
  float4 g = make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
  if (rep.x >= 0) {
    const OrdGeom og = a.ordg[col];
    double xe = a.g.gridpos[rep.x * 3], ye = a.g.gridpos[rep.x * 3 + 1], ze = a.g.gridpos[rep.x * 3 + 2];
    TraceStep t;
    if (trace_step(a.g, og, rep.y, xe, ye, ze, t))
      g = make_float4((float)t.u, (float)t.v, (float)t.so, __int_as_float(t.face));
  }
  const_cast<float4*>(a.geo)[i] = g;
}

// Once per solve: the first cell step of every point's characteristic for every
// ordinate (the loop body of BACK_INT_GRID3D_UNPOL, shdomsub1.f:3822-3920), in
// the ordinate's ring order.  A step whose face corners are known (boundary
// points or earlier points of the sweep order), lie within the last
// RING_WINDOW levels and whose face, u, v and path length match its class's
// becomes a one-cell record: the four corners' ring slots and the extinction
// on the face (the cell integral is evaluated in the sweep, from the class's
// path length and the point's extinction).  Every other point is marked for
// the general trace and its level flagged; boundary entries carry their index.
//
// A wavefront takes one entry of one octant for up to 64 of the octant's
// ordinates (lane = ordinate), so the point's geometry, ranks and levels are
// read once per wavefront and each lane writes one 16-byte record.
__device__ inline void sweep_record(const SweepArgs& a, int o, int col, int jl);

// Which far faces of each point's start cell a one-cell step may end on: all
// four corners computed before the point in the reference's sweep order
// (preset boundary points or lower rank) and held by the ring (at most
// RING_WINDOW levels back).  The same for every ordinate of an octant, so
// made once per (octant, point) rather than per record.
__global__ __launch_bounds__(256) void k_face_mask(SweepArgs a) {
  const int o = blockIdx.y;
  const int jl = blockIdx.x * blockDim.x + threadIdx.x;
  if (a.ring_nlev[o] <= 0 || jl >= a.npts) return;
  const size_t ob = (size_t)o * a.npts;
  const int code = a.ring_code[ob + jl];
  if (code < 0) return;   // boundary entry
  const int pt = code, cell = a.ring_cell[ob + jl];
  const int bitx = o & 1, bity = (o >> 1) & 1, bitz = (o >> 2) & 1;
  const int presetbit = bitz ? 2 : 1;
  const int32_t* rank = a.rank + ob;
  const int32_t* rlev = a.ring_level + ob;
  const int L = rlev[pt], rp = rank[pt];
  uint32_t m = 0;
  for (int d = 0; d < 3; d++) {
    const int face = d == 0 ? 1 - bitx : d == 1 ? 3 - bity : 5 - bitz;
    bool ok = true;
    for (int k = 0; k < 4 && ok; k++) {
      const int q = a.g.gridptr[(size_t)cell * 8 + face_corner(face, k)];
      const int lq = rlev[q];
      ok = ((a.preset[q] & presetbit) || rank[q] < rp) && lq < L && lq >= L - RING_WINDOW;
    }
    if (ok) m |= 1u << d;
  }
  const_cast<uint8_t*>(a.face_ok)[ob + pt] = (uint8_t)m;
}

__global__ __launch_bounds__(256) void k_sweep_records(SweepArgs a) {
  const int o = blockIdx.y;
  const int jl = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (a.ring_nlev[o] <= 0 || jl >= a.npts) return;
  for (int k = threadIdx.x & 63; k < a.oct_nord[o]; k += 64) sweep_record(a, o, a.oct_col0[o] + k, jl);
}

__device__ inline void sweep_record(const SweepArgs& a, int o, int col, int jl) {
#pragma clang fp contract(off)
  const size_t ob = (size_t)o * a.npts;
  const OrdGeom og = a.ordg[col];
  const int code = a.ring_code[ob + jl];
  const int32_t* pos = a.oct_pos + ob;
  SweepRec r;
  r.r[0] = r.r[1] = r.r[2] = r.r[3] = 0;
  r.ext0 = 0.0f;
  if (code < 0) {
    // boundary entry: its radiance is the boundary condition
    const int b = -1 - code;
    const int pt = (og.bitz ? a.bcbot : a.bctop)[b];
    r.head = (uint32_t)pt | ((uint32_t)a.ring_staged[ob + pt] << 29) | (2u << 30);
    r.r[0] = (uint16_t)(b & 0xffff);
    r.r[1] = (uint16_t)(b >> 16);
    // the radiance itself in r[2..3]: the sky radiance of a downwelling
    // ordinate here; the surface radiance of an upwelling one is written by
    // k_bottom_bc every iteration
    const uint32_t bits = __float_as_uint(og.bitz ? 0.0f : a.skyrad[col]);
    r.r[2] = (uint16_t)(bits & 0xffff);
    r.r[3] = (uint16_t)(bits >> 16);
  } else {
    const int pt = code;
    const int cell = a.ring_cell[ob + jl];
    const int8_t fl = a.g.cellflags[cell] & 3;
    const uint32_t fm = a.face_ok[ob + pt];
    double xe = a.g.gridpos[pt * 3], ye = a.g.gridpos[pt * 3 + 1], ze = a.g.gridpos[pt * 3 + 2];
    TraceStep t;
    bool ok = trace_step(a.g, og, cell, xe, ye, ze, t) && ((fm >> (t.face >> 1)) & 1);
    if (ok)
      for (int k = 0; k < 4; k++) r.r[k] = (uint16_t)(pos[t.i[k]] % a.slots);
    const float4 g = a.geo[(size_t)col * a.ncls + step_class(a, pt, fl)];
    const int gface = __float_as_int(g.w);
    int cn[4] = {t.i[0], t.i[1], t.i[2], t.i[3]};
    double u = t.u, v = t.v;
    if (ok && t.so > 0.0 && gface >= 0 && gface != t.face) {
      // a step through an edge shared with the class's face: on that face
      face_uv(a.g, cell, gface, xe, ye, ze, u, v, cn);
      ok = (fm >> (gface >> 1)) & 1;
      if (ok)
        for (int k = 0; k < 4; k++) r.r[k] = (uint16_t)(pos[cn[k]] % a.slots);
    }
    // the step is its class's when u, v and the path length agree to within
    // the rounding of the float grid positions
    const double tol = a.cls_tol;
    const bool cls_ok = gface >= 0 && fabs(u - (double)g.x) <= tol && fabs(v - (double)g.y) <= tol &&
                        fabs(t.so - (double)g.z) <= tol * fabs(t.so);
    // a zero-length step ending on a corner (seams of periodic grids) takes
    // that corner's radiance
    int corner = -1;
    if (ok && t.so == 0.0 && (t.u == 0.0 || t.u == 1.0) && (t.v == 0.0 || t.v == 1.0))
      corner = (t.u == 1.0 ? 2 : 0) + (t.v == 1.0 ? 1 : 0);
    const uint32_t staged = (uint32_t)a.ring_staged[ob + pt];
    if (ok && corner >= 0) {
      // (a select chain: a runtime index would put the record in scratch)
      r.r[0] = corner == 0 ? r.r[0] : corner == 1 ? r.r[1] : corner == 2 ? r.r[2] : r.r[3];
      r.head = (uint32_t)pt | ((uint32_t)fl << RECORD_PT_BITS) | (staged << 29) | (3u << 30);
    } else if (ok && cls_ok) {
      const float* ext = a.ext;
      const double uc = g.x, vc = g.y;
      const double f1 = (1 - uc) * (1 - vc), f2 = (1 - uc) * vc;
      const double f3 = uc * (1 - vc), f4 = uc * vc;
      r.ext0 = (float)(f1 * ext[cn[0]] + f2 * ext[cn[1]] + f3 * ext[cn[2]] + f4 * ext[cn[3]]);
      r.head = (uint32_t)pt | ((uint32_t)fl << RECORD_PT_BITS) | (staged << 29);
    } else {
      r.head = (uint32_t)pt | ((uint32_t)fl << RECORD_PT_BITS) | (staged << 29) | (1u << 30);
      atomicAdd(a.ngeneral, 1);
      atomicOr(a.col_general + col, 1);
      const_cast<uint8_t*>(a.gflag)[(size_t)col * a.gld + a.ring_level[ob + pt]] = 1;
    }
  }
  const_cast<uint4*>(a.rec)[(size_t)col * a.npts + jl] = *reinterpret_cast<const uint4*>(&r);
}

// Barrier that orders LDS traffic only: outstanding global loads (the next
// levels' records) and stores (radiances) stay in flight across it.
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}


// The cell integral of a one-cell step (shdomsub1.f:3948-3987, with the
// reference's tau branches) as coefficients of the face and point sources:
// src x abs = b0 SRCEXT0 + b1 SRCEXT1 before the clamp at zero, and the
// cell's transmission, in single precision from the face and point
// extinctions and the path length.
__device__ inline void cell_coefs(float ext0, float ext1, float so, float& b0, float& b1, float& trc) {
  const float e = 0.5f * (ext0 + ext1);
  const float tau = e * so;
  float absc;
  if (tau >= 0.5f) {
    trc = __expf(-tau);
    absc = 1.0f - trc;
  } else {
    absc = tau * (1.0f - 0.5f * tau * (1.0f - 0.33333333333f * tau * (1.0f - 0.25f * tau)));
    trc = 1.0f - absc;
  }
  float c0, c1;
  if (tau <= 2.0f) {
    if (e == 0.0f) {
      c0 = c1 = 0.0f;
    } else {
      const float k = 0.08333333333f * so, re = __builtin_amdgcn_rcpf(e);
      c0 = (0.5f - k * ext1) * re;
      c1 = (0.5f + k * ext0) * re;
    }
  } else {
    float e0p = ext0, sc = 1.0f;
    if (tau > 4.0f) {
      e0p = ext1 + (ext0 - ext1) * 4.0f / tau;
      if (ext0 > 0.0f) sc = e0p / ext0;
    }
    const float d = e0p + ext1;
    const float g = 2.0f / d * (1.0f - 2.0f / tau + 2.0f * trc / absc);
    c0 = sc * (1.0f - ext1 * g) / d;
    c1 = (1.0f + e0p * g) / d;
  }
  b0 = c0 * absc;
  b1 = c1 * absc;
}

constexpr uint32_t RECORD_PT_MASK = (1u << RECORD_PT_BITS) - 1;

struct RingSet {
  uint4 w0[RING_K];
  float s[RING_K];   // the entry's source x extinction
  float e[RING_K];   // and its extinction
  uint4 sec;     // ring slots of the sector this thread writes after the level
  int sec_p0;    // its first point
  bool sec_ok;   // there is such a sector (known from indices, not from the loads)
};

// One ordinate (workgroup) through every level of its octant.  LDS holds a
// ring of (radiance, source) pairs of the last levels (a level reads
// RING_WINDOW levels back; a sector written after a level spans up to eight
// levels), then the level offsets, the sector offsets and the full-barrier
// flags.  Radiances go to the point-ordered row one sector (32 or 16 bytes)
// at a time, after the sector's last point, except in columns with
// multi-cell traces (which read earlier radiances from the row) and for
// points of sectors that span more levels.
//
// Two instantiations: FULL handles every case (multi-cell traces, levels with
// more entries than the threads' register slots, more sectors than threads);
// the fast one has none of those paths, so its level loop carries nothing
// but the record prefetch, and runs the columns (and grids) that need none
// of them.  Each launch covers a hemisphere's columns; a workgroup whose
// column belongs to the other instantiation returns at once.
//
// R4: the ring holds radiances alone (4 bytes a slot, for grids whose
// (radiance, source) ring does not fit in LDS); a face corner's source is
// read from the column's ring-ordered source row at the corner's position,
// which its slot gives (the ring holds the last C positions).
template <bool FULL, bool R4>
__global__ __launch_bounds__(RING_THREADS) void k_sweep_ring(SweepArgs a, int col0) {
  if (a.stop && *a.stop) return;
  extern __shared__ float4 lds4[];
  const int col = col0 + blockIdx.x;
  if (FULL != (a.ring_full || a.col_general[col] != 0)) return;
  if (blockIdx.y == 0 && col >= a.qu_only_from) return;
  // Stokes component of this workgroup: Q and U (kst 1, 2) see zero boundary
  // radiances and keep the sign of the cell source (BACK_INT_GRID3D clamps
  // only I, shdomsub1.f:3640)
  const int kst = blockIdx.y;
  const int o = a.col_oct[col];
  const int nlev = a.ring_nlev[o];
  const OrdGeom og = a.ordg[col];
  const int up = og.bitz;
  const int C = a.slots;
  // the step classes: in LDS for the fast instantiation, read from device
  // memory (a few KB per column, cache-resident) by the full one, whose rings
  // may need all of the LDS
  float4* geo_lds = lds4;                                   // [ncls] (fast only)
  const float4* geo = FULL ? a.geo + (size_t)col * a.ncls : geo_lds;
  float2* ring = reinterpret_cast<float2*>(lds4 + (FULL ? 0 : a.ncls));
  float* ring1 = reinterpret_cast<float*>(ring);   // R4
  int32_t* off = R4 ? reinterpret_cast<int32_t*>(ring1 + C) : reinterpret_cast<int32_t*>(ring + C);
  int32_t* soffl = off + a.gld;                     // [gld + 1] sectors by completion level
  uint8_t* full = reinterpret_cast<uint8_t*>(soffl + a.gld + 1);
  const int32_t* goff = a.ring_off + (size_t)o * a.gld;
  const float* se = a.srcext + kst * a.pdo + (size_t)col * a.ldp;  // ring order
  float* rad = a.rad + kst * a.pdo + (size_t)col * a.ldp;          // point order
  const uint4* rec0 = a.rec + (size_t)col * a.npts;
  const float* ep = a.ext_pos + (size_t)o * a.ldp;  // ring order
  const uint8_t* gf = a.gflag + (size_t)col * a.gld;
  const int32_t* pos = a.oct_pos + (size_t)o * a.npts;
  const int32_t* soff = a.sec_off + (size_t)o * (a.gld + 1);
  const int32_t* sp0 = a.sec_p0 + a.sec_base[o];
  const uint4* sslot = a.sec_slot + a.sec_base[o];
  // a ring slot's (radiance, source): with R4 the source at the slot's
  // position in [lo, lo + C), lo = the level's end - C, lom = lo mod C
  auto gather = [&](uint32_t slot, int lo, int lom) -> float2 {
    if (!R4) return ring[slot];
    int d = (int)slot - lom;
    if (d < 0) d += C;
    return make_float2(ring1[slot], se[lo + d]);
  };
  auto rad_of = [&](uint32_t slot) { return R4 ? ring1[slot] : ring[slot].x; };
  auto lo_mod = [&](int lo) {
    int m = lo % C;
    return m < 0 ? m + C : m;
  };
  const bool wide = a.sector_points == 8;
  const int tid = threadIdx.x;
  const float sky = kst ? 0.0f : a.skyrad[col];
  const bool staging = a.col_general[col] == 0;
  const int nbc = up ? a.nbot : a.ntop;
  const int32_t* bc = up ? a.bcbot : a.bctop;
  // boundary radiances (PATH_INTEGRATION shdomsub1.f:2042-2086) into the
  // point-ordered row first: multi-cell traces read them there
  for (int b = tid; b < nbc; b += RING_THREADS) rad[bc[b]] = kst ? 0.0f : (up ? a.bcrad_bot[b] : sky);
  for (int L = tid; L <= nlev; L += RING_THREADS) off[L] = goff[L];
  for (int L = tid; L <= nlev + 1; L += RING_THREADS) soffl[L] = soff[L];
  for (int L = tid; L < a.gld; L += RING_THREADS) full[L] = gf[L];
  if (!FULL)
    for (int i = tid; i < a.ncls; i += RING_THREADS) geo_lds[i] = a.geo[(size_t)col * a.ncls + i];
  const int nsw = goff[nlev];
  // a one-cell step's radiance from its face corners, its class and the
  // point's source and extinction
  auto one_cell = [&](const uint4& w, const float2& q0, const float2& q1, const float2& q2,
                      const float2& q3, const float4& g, float s1, float e1) {
    const float u = g.x, v = g.y;
    const float f1 = (1.f - u) * (1.f - v), f2 = (1.f - u) * v, f3 = u * (1.f - v), f4 = u * v;
    const float s0 = f1 * q0.y + f2 * q1.y + f3 * q2.y + f4 * q3.y;
    const float r0 = f1 * q0.x + f2 * q1.x + f3 * q2.x + f4 * q3.x;
    float b0, b1, trc;
    cell_coefs(__uint_as_float(w.w), e1, g.z, b0, b1, trc);
    const float sv = b0 * s0 + b1 * s1;
    return (kst ? sv : fmaxf(0.f, sv)) + trc * r0;
  };
  auto class_of = [&](uint32_t head) {
    return step_class(a, (int)(head & RECORD_PT_MASK), (int8_t)((head >> RECORD_PT_BITS) & 3));
  };
  auto put = [&](int jl, int k0, float val, float s1) {
    int i = k0 + jl;
    if (i >= C) i -= C;
    if (R4)
      ring1[i] = val;
    else
      ring[i] = make_float2(val, s1);
  };
  // Entries past the register slots (levels longer than RING_K * RING_THREADS,
  // full instantiation only), RING_KX per thread at a time: their records,
  // sources and extinctions are all loaded, then all LDS gathers issued,
  // then the cell integrals, so one memory latency covers RING_KX entries.
  // Multi-cell records (kind 1) are left to the caller's trace pass.
  auto overflow = [&](int j0, int end, int k0) {
    uint4 w[RING_KX];
    float sx[RING_KX], ex[RING_KX];
#pragma unroll
    for (int u = 0; u < RING_KX; u++) {
      const int jl = j0 + u * RING_THREADS;
      w[u] = make_uint4(1u << 30, 0u, 0u, 0u);
      sx[u] = ex[u] = 0.f;
      if (jl < end) {
        w[u] = rec0[jl];
        sx[u] = se[jl];
        ex[u] = ep[jl];
      }
    }
    float2 q[RING_KX][4];
    float4 gq[RING_KX];
    const int lo = end - C, lom = lo_mod(lo);
#pragma unroll
    for (int u = 0; u < RING_KX; u++) {
      const int kind = w[u].x >> 30;
      if (kind == 0 || kind == 3) {
        q[u][0] = gather(w[u].y & 0xffffu, lo, lom);
        q[u][1] = gather(w[u].y >> 16, lo, lom);
        q[u][2] = gather(w[u].z & 0xffffu, lo, lom);
        q[u][3] = gather(w[u].z >> 16, lo, lom);
      }
      if (kind == 0) gq[u] = geo[class_of(w[u].x)];
    }
#pragma unroll
    for (int u = 0; u < RING_KX; u++) {
      const uint32_t head = w[u].x;
      const int kind = head >> 30;
      if (kind == 1) continue;
      float val;
      if (kind == 2) {
        val = kst ? 0.0f : __uint_as_float(w[u].z);
      } else {
        val = kind == 3 ? q[u][0].x : one_cell(w[u], q[u][0], q[u][1], q[u][2], q[u][3], gq[u], sx[u], ex[u]);
        if (!(staging && ((head >> 29) & 1))) rad[head & RECORD_PT_MASK] = val;
      }
      put(j0 + u * RING_THREADS, k0, val, sx[u]);
    }
  };
  // Records of a level: register slot k of a wave is loaded only when the
  // wave's first entry of that slot lies in the level (a wave-uniform
  // branch); lanes past the level's end load its last record, one cache line
  // for the wave.  With them the sectors that complete on the level before
  // (written after level L), loaded by the waves that have one.
  const int wfirst = tid & ~63;
  auto load = [&](int L, RingSet& S) {
    const int beg = __builtin_amdgcn_readfirstlane(off[min(L - 1, nlev)]);
    const int end = __builtin_amdgcn_readfirstlane(off[min(L, nlev)]);
    const int last = max(end - 1, 0);
#pragma unroll
    for (int k = 0; k < RING_K; k++) {
      if (beg + wfirst + k * RING_THREADS < end) {
        const int jl = min(beg + tid + k * RING_THREADS, last);
        S.w0[k] = rec0[jl];
        S.s[k] = se[jl];
        S.e[k] = ep[jl];
      }
    }
    S.sec_ok = false;
    if (staging && L >= 2 && L <= nlev + 1) {
      const int s0 = soffl[L - 1], s1 = soffl[L];
      S.sec_ok = s0 + tid < s1;
      if (__builtin_amdgcn_readfirstlane(s0 + wfirst) < s1) {
        const int sct = min(s0 + tid, s1 - 1);  // clamped, used only if sec_ok
        S.sec = sslot[sct];
        S.sec_p0 = sp0[sct];
      }
    }
  };
  auto write_sector = [&](uint4 sl, int p0) {
    float4 v;
    v.x = rad_of(sl.x & 0xffffu);
    v.y = rad_of(sl.x >> 16);
    v.z = rad_of(sl.y & 0xffffu);
    v.w = rad_of(sl.y >> 16);
    *reinterpret_cast<float4*>(rad + p0) = v;
    if (wide) {
      v.x = rad_of(sl.z & 0xffffu);
      v.y = rad_of(sl.z >> 16);
      v.z = rad_of(sl.w & 0xffffu);
      v.w = rad_of(sl.w >> 16);
      *reinterpret_cast<float4*>(rad + p0 + 4) = v;
    }
  };
  // the whole sectors whose last point is on level L from index `first` on
  // (those beyond one per thread, and those of the last level), not prefetched
  auto flush_rest = [&](int L, int first) {
    const int s1 = soffl[L + 1];
    for (int sct = soffl[L] + first + tid; sct < s1; sct += RING_THREADS) write_sector(sslot[sct], sp0[sct]);
  };
  // Level L from `cur` while level L+1 streams into `fut`.  First every
  // register of `cur` is taken as read: one wait for all of the loads
  // outstanding, which are this level's records (issued a level ago) and the
  // stores of the level before; the next level's loads, issued after it, are
  // never waited for here, and may be conditional without widening any wait.
  // All of the level's LDS reads (face corners: slots of levels L-3 .. L-1;
  // the sector completed on level L-1: levels L-4 .. L-1) come before any of
  // its LDS writes (slots of level L); they are issued before the next
  // level's loads and waited for once, after them.
  auto step = [&](int L, RingSet& cur, RingSet& fut) {
#pragma unroll
    for (int k = 0; k < RING_K; k++)
      asm volatile("" ::"v"(cur.w0[k].x), "v"(cur.w0[k].y), "v"(cur.w0[k].z), "v"(cur.w0[k].w),
                   "v"(cur.s[k]), "v"(cur.e[k]));
    asm volatile("" ::"v"(cur.sec.x), "v"(cur.sec.y), "v"(cur.sec.z), "v"(cur.sec.w), "v"(cur.sec_p0));
    const int beg = __builtin_amdgcn_readfirstlane(off[L - 1]);
    const int end = __builtin_amdgcn_readfirstlane(off[L]);
    const int k0 = beg % C - beg;  // slot of position jl: k0 + jl, wrapped once
    const int lo = end - C, lom = lo_mod(lo);
    float2 g[RING_K][4];
    float4 gc[RING_K];   // the steps' classes
#pragma unroll
    for (int k = 0; k < RING_K; k++) {
      const int jl = beg + tid + k * RING_THREADS;
      const uint4 w0 = cur.w0[k];
      if (jl < end && (w0.x >> 30) != 1 && (w0.x >> 30) != 2) {
        g[k][0] = gather(w0.y & 0xffffu, lo, lom);
        g[k][1] = gather(w0.y >> 16, lo, lom);
        g[k][2] = gather(w0.z & 0xffffu, lo, lom);
        g[k][3] = gather(w0.z >> 16, lo, lom);
        gc[k] = geo[class_of(w0.x)];
      }
    }
    const bool dosec = staging && L > 1 && cur.sec_ok;
    float4 sv = make_float4(0.f, 0.f, 0.f, 0.f), sw = sv;
    if (dosec) {
      const uint4 sl = cur.sec;
      sv.x = rad_of(sl.x & 0xffffu);
      sv.y = rad_of(sl.x >> 16);
      sv.z = rad_of(sl.y & 0xffffu);
      sv.w = rad_of(sl.y >> 16);
      if (wide) {
        sw.x = rad_of(sl.z & 0xffffu);
        sw.y = rad_of(sl.z >> 16);
        sw.z = rad_of(sl.w & 0xffffu);
        sw.w = rad_of(sl.w >> 16);
      }
    }
    load(L + 1, fut);  // past the last level this loads nothing
    uint32_t gen = 0;  // this thread's multi-cell traces among its first RING_K entries
#pragma unroll
    for (int k = 0; k < RING_K; k++) {
      const int jl = beg + tid + k * RING_THREADS;
      if (jl >= end) continue;
      const uint4 w0 = cur.w0[k];
      const uint32_t head = w0.x;
      const int kind = head >> 30;
      if (kind == 1) {
        gen |= 1u << k;
        continue;
      }
      float val;
      if (kind == 2) {
        val = kst ? 0.0f : __uint_as_float(w0.z);
      } else {
        // kind 3: a zero-length step onto a corner, whose radiance it takes
        val = kind == 3 ? g[k][0].x
                        : one_cell(w0, g[k][0], g[k][1], g[k][2], g[k][3], gc[k], cur.s[k], cur.e[k]);
        if (!(staging && ((head >> 29) & 1))) rad[head & RECORD_PT_MASK] = val;
      }
      put(jl, k0, val, cur.s[k]);
    }
    if (dosec) {
      *reinterpret_cast<float4*>(rad + cur.sec_p0) = sv;
      if (wide) *reinterpret_cast<float4*>(rad + cur.sec_p0 + 4) = sw;
    }
    if (FULL)
      for (int j0 = beg + tid + RING_K * RING_THREADS; j0 < end; j0 += RING_KX * RING_THREADS)
        overflow(j0, end, k0);
    if (FULL && __builtin_amdgcn_readfirstlane(full[L])) {
      // multi-cell traces read radiances of earlier levels from the row,
      // complete and visible after the full barrier that ended level L-1
      int k = 0;
      for (int jl = beg + tid; jl < end; jl += RING_THREADS, k++) {
        const uint32_t head = rec0[jl].x;
        const bool gn = k < RING_K ? ((gen >> k) & 1) != 0 : (head >> 30) == 1;
        if (!gn) continue;
        const int p = head & RECORD_PT_MASK;
        const float s1 = se[jl];
        const float val = sweep_general(a, og, col, o, p, a.ring_cell[(size_t)o * a.npts + jl], L, se,
                                        pos, kst);
        rad[p] = val;
        put(jl, k0, val, s1);
      }
    }
    if (FULL && staging && L > 1 && soffl[L] - soffl[L - 1] > RING_THREADS) flush_rest(L - 1, RING_THREADS);
    if (FULL && L < nlev && __builtin_amdgcn_readfirstlane(full[L + 1]))
      __syncthreads();
    else
      lds_barrier();
  };
  RingSet S0, S1;
  __syncthreads();
  load(1, S0);
  int L = 1;
  for (; L + 1 <= nlev; L += 2) {
    step(L, S0, S1);
    step(L + 1, S1, S0);
  }
  if (L == nlev) step(L, S0, S1);
  if (staging && nlev >= 1) flush_rest(nlev, 0);
}

// Boundary radiances for one hemisphere: sky radiance at the top points for
// downwelling ordinates, surface radiance at the bottom points for upwelling
// ones (PATH_INTEGRATION shdomsub1.f:2042-2086).
// blockIdx.y: the Stokes component (Q and U boundary radiances are zero).
__global__ void k_sweep_preset(SweepArgs a, int up, int col0, int ncol) {
  if (a.stop && *a.stop) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int np = up ? a.nbot : a.ntop;
  if (i >= np * ncol) return;
  const int b = i / ncol, c = col0 + (i - b * ncol);
  const int kst = blockIdx.y;
  float* rad = a.rad + kst * a.pdo;
  if (up)
    rad[(size_t)c * a.ldp + a.bcbot[b]] = kst ? 0.0f : a.bcrad_bot[b];
  else
    rad[(size_t)c * a.ldp + a.bctop[b]] = kst ? 0.0f : a.skyrad[c];
}

// FIXED_LAMBERTIAN_BOUNDARY (shdomsub1.f:2438-2480) with the downwelling
// hemispheric flux of each bottom point (:2118-2135): one wavefront per point,
// lane l summing the ordinates l, l+64, ... in reference order, then a fixed
// butterfly over the lanes (deterministic; the reference sums serially).
__global__ __launch_bounds__(256) void k_bottom_bc(const float* __restrict__ rad, int ldp,
                                                   const int32_t* __restrict__ bcbot, int nbot,
                                                   const int32_t* __restrict__ col_of_ref,
                                                   const float* __restrict__ fluxw, int ndown,
                                                   const float* __restrict__ dirflux, float albpi,
                                                   float* __restrict__ bcrad, BottomRecs br,
                                                   const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int b = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (b >= nbot) return;
  const int p = bcbot[b];
  float f = 0.f;
  for (int k = lane; k < ndown; k += 64) f = f + fluxw[k] * rad[(size_t)col_of_ref[k] * ldp + p];
  for (int m = 32; m >= 1; m >>= 1) f = f + __shfl_xor(f, m);
  const float v = albpi * (dirflux[p] + f);
  if (lane == 0) bcrad[b] = v;
  if (br.rec) {
    // the surface radiance into the boundary entry (r[2..3], the third word)
    // of every upwelling column's records
    for (int c = br.col0 + lane; c < br.col1; c += 64) {
      const int o = br.col_oct[c];
      const size_t e = (size_t)c * br.npts + br.oct_pos[(size_t)o * br.npts + p];
      reinterpret_cast<uint32_t*>(br.rec + e)[2] = __float_as_uint(v);
    }
  }
}

// Hemispheric fluxes of every point, summed serially in reference order with
// the loads of eight ordinates in flight.
__global__ void k_fluxes(const float* __restrict__ rad, int ldp, const int32_t* __restrict__ col_of_ref,
                         const float* __restrict__ fluxw, int ndown, int nang, int npts,
                         float* __restrict__ fluxes) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  auto hemisphere = [&](int k0, int k1) {
    float s = 0.f;
    int k = k0;
    for (; k + 8 <= k1; k += 8) {
      float v[8];
#pragma unroll
      for (int q = 0; q < 8; q++) v[q] = rad[(size_t)col_of_ref[k + q] * ldp + p];
#pragma unroll
      for (int q = 0; q < 8; q++) s = s + fluxw[k + q] * v[q];
    }
    for (; k < k1; k++) s = s + fluxw[k] * rad[(size_t)col_of_ref[k] * ldp + p];
    return s;
  };
  fluxes[p] = hemisphere(0, ndown);
  fluxes[npts + p] = hemisphere(ndown, nang);
}

void launch_sweep_preset(hipStream_t st, const SweepArgs& a, bool up, int col0, int ncol) {
  const int n = (up ? a.nbot : a.ntop) * ncol;
  if (n <= 0) return;
  k_sweep_preset<<<dim3((n + 255) / 256, a.nst), 256, 0, st>>>(a, up ? 1 : 0, col0, ncol);
}

void launch_sweep_columns(hipStream_t st, const SweepArgs& a, int col0, int ncol) {
  if (ncol <= 0) return;
  k_sweep_columns<<<dim3(ncol, a.nst), SWEEP_THREADS, 0, st>>>(a, col0);
}

void launch_face_mask(hipStream_t st, const SweepArgs& a) {
  k_face_mask<<<dim3((a.npts + 255) / 256, 8), 256, 0, st>>>(a);
}

void launch_sweep_records(hipStream_t st, const SweepArgs& a) {
  k_sweep_records<<<dim3((a.npts + 3) / 4, 8), 256, 0, st>>>(a);
}

void launch_sweep_geo(hipStream_t st, const SweepArgs& a, int ncols) {
  const int n = ncols * a.ncls;
  if (n > 0) k_sweep_geo<<<(n + 255) / 256, 256, 0, st>>>(a, ncols);
}

void launch_sweep_ring(hipStream_t st, const SweepArgs& a, int col0, int ncol, size_t lds_fast,
                       size_t lds_full, bool fast, bool full) {
  if (ncol <= 0) return;
  const dim3 grid(ncol, a.nst);
  if (a.ring4) {
    if (fast) k_sweep_ring<false, true><<<grid, RING_THREADS, lds_fast, st>>>(a, col0);
    if (full) k_sweep_ring<true, true><<<grid, RING_THREADS, lds_full, st>>>(a, col0);
    return;
  }
  if (fast) k_sweep_ring<false, false><<<grid, RING_THREADS, lds_fast, st>>>(a, col0);
  if (full) k_sweep_ring<true, false><<<grid, RING_THREADS, lds_full, st>>>(a, col0);
}

int sweep_ring_max_lds(size_t lds_fast, size_t lds_full, bool ring4) {
  const void* kf = ring4 ? reinterpret_cast<const void*>(k_sweep_ring<false, true>)
                         : reinterpret_cast<const void*>(k_sweep_ring<false, false>);
  const void* ku = ring4 ? reinterpret_cast<const void*>(k_sweep_ring<true, true>)
                         : reinterpret_cast<const void*>(k_sweep_ring<true, false>);
  const bool ok =
      (lds_fast == 0 ||
       hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_fast) == hipSuccess) &&
      hipFuncSetAttribute(ku, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_full) == hipSuccess;
  return ok ? 0 : 1;
}

void launch_bottom_bc(hipStream_t st, const float* rad, int ldp, const int32_t* bcbot, int nbot,
                      const int32_t* col_of_ref, const float* fluxw, int ndown,
                      const float* dirflux, float albpi, float* bcrad, const BottomRecs& br,
                      const int* stop) {
  if (nbot <= 0) return;
  k_bottom_bc<<<(nbot + 3) / 4, 256, 0, st>>>(rad, ldp, bcbot, nbot, col_of_ref, fluxw, ndown, dirflux,
                                               albpi, bcrad, br, stop);
}

void launch_fluxes(hipStream_t st, const float* rad, int ldp, const int32_t* col_of_ref,
                   const float* fluxw, int ndown, int nang, int npts, float* fluxes) {
  k_fluxes<<<(npts + 63) / 64, 64, 0, st>>>(rad, ldp, col_of_ref, fluxw, ndown, nang, npts, fluxes);
}

}  // namespace at3d
