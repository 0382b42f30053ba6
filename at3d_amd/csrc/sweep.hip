// Discrete-ordinate sweeps (BACK_INT_GRID3D_UNPOL, shdomsub1.f:3695-4032).
//
// The reference integrates one ordinate at a time, visiting grid points in its
// sweep order (SWEEPING_ORDER :3261) and tracing each point back to the first
// cell face whose four corner radiances are already known.  The host planner
// replays that order and assigns every point a dependency level per octant
// (all upwind face points of every ordinate of the octant sit on lower
// levels).  On the GPU one workgroup integrates one ordinate over the whole
// grid: it walks the levels of the ordinate's octant in order, a thread per
// point of the level, with a workgroup barrier between levels, so a
// hemisphere's sweep is one launch whose workgroups run independently.  The
// DO arrays are stored ordinate-major ([ordinate][point]) so a workgroup's
// gathers and stores stay inside one contiguous row.
#include <hip/hip_runtime.h>
#include <limits.h>

#include "kernels.h"

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
                                           SimpleStep& st) {
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
  const float* se = a.srcext + (size_t)col * a.ldp;
  const double f1 = (1 - u) * (1 - v), f2 = (1 - u) * v, f3 = u * (1 - v), f4 = u * v;
  const double ext0 = f1 * ext[i1] + f2 * ext[i2] + f3 * ext[i3] + f4 * ext[i4];
  const double srcext0 = f1 * se[i1] + f2 * se[i2] + f3 * se[i3] + f4 * se[i4];
  double src, absc, trc;
  cell_integral(ext0, (double)ext[pt], srcext0, (double)se[pt], so, src, absc, trc);
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

__device__ inline float sweep_simple(const SweepArgs& a, const OrdGeom& og, int col, int pt) {
  SimpleStep st;
  sweep_simple_static(a, og, col, pt, st);
  return sweep_simple_finish(st, a.rad + (size_t)col * a.ldp);
}

// General characteristic: trace through cells until a face with known radiances.
__device__ inline float sweep_general(const SweepArgs& a, const OrdGeom& og, int col, int o, int pt,
                                      int icell, int L) {
  const int presetbit = og.bitz ? 2 : 1;
  const int32_t* rank = a.rank + (size_t)o * a.npts;
  const int32_t* level = a.level + (size_t)o * a.npts;
  const int rp = rank[pt];
  const float* ext = a.ext;
  const float* se = a.srcext + (size_t)col * a.ldp;
  const float* rad = a.rad + (size_t)col * a.ldp;
  double xe = a.g.gridpos[pt * 3], ye = a.g.gridpos[pt * 3 + 1], ze = a.g.gridpos[pt * 3 + 2];
  double ext1 = ext[pt], srcext1 = se[pt];
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
    const double srcext0 = f1 * se[t.i[0]] + f2 * se[t.i[1]] + f3 * se[t.i[2]] + f4 * se[t.i[3]];
    double src, absc, trc;
    cell_integral(ext0, ext1, srcext0, srcext1, t.so, src, absc, trc);
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
  const int col = col0 + blockIdx.x;
  const int o = a.col_oct[col];
  const OrdGeom og = a.ordg[col];
  const int nlev = a.nlev[o];
  const int32_t* off = a.lvl_off + a.lo_base[o];
  float* rad = a.rad + (size_t)col * a.ldp;
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
      if (code[k] >= 0 && !general) sweep_simple_static(a, og, col, code[k], st[k]);
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
        rad[pt] = sweep_general(a, og, col, o, pt, a.lvl_cell[j], L);
      }
    }
    for (int j = beg + tid + SWEEP_K * SWEEP_THREADS; j < end; j += SWEEP_THREADS) {
      const int c = a.lvl_pts[j];
      if (c >= 0 && !general) {
        rad[c] = sweep_simple(a, og, col, c);
      } else {
        const int pt = c >= 0 ? c : ~c;
        rad[pt] = sweep_general(a, og, col, o, pt, a.lvl_cell[j], L);
      }
    }
    if (L < nlev) prepare(L + 1);
  }
}

// Boundary radiances for one hemisphere: sky radiance at the top points for
// downwelling ordinates, surface radiance at the bottom points for upwelling
// ones (PATH_INTEGRATION shdomsub1.f:2042-2086).
__global__ void k_sweep_preset(SweepArgs a, int up, int col0, int ncol) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int np = up ? a.nbot : a.ntop;
  if (i >= np * ncol) return;
  const int b = i / ncol, c = col0 + (i - b * ncol);
  if (up)
    a.rad[(size_t)c * a.ldp + a.bcbot[b]] = a.bcrad_bot[b];
  else
    a.rad[(size_t)c * a.ldp + a.bctop[b]] = a.skyrad[c];
}

// FIXED_LAMBERTIAN_BOUNDARY (shdomsub1.f:2438-2480) with the downwelling
// hemispheric flux summed in the reference's ordinate order (:2118-2135).
__global__ void k_bottom_bc(const float* __restrict__ rad, int ldp, const int32_t* __restrict__ bcbot,
                            int nbot, const int32_t* __restrict__ col_of_ref,
                            const float* __restrict__ fluxw, int ndown,
                            const float* __restrict__ dirflux, float albpi,
                            float* __restrict__ bcrad) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbot) return;
  const int p = bcbot[b];
  float f = 0.f;
  for (int k = 0; k < ndown; k++) f = f + fluxw[k] * rad[(size_t)col_of_ref[k] * ldp + p];
  bcrad[b] = albpi * (dirflux[p] + f);
}

__global__ void k_fluxes(const float* __restrict__ rad, int ldp, const int32_t* __restrict__ col_of_ref,
                         const float* __restrict__ fluxw, int ndown, int nang, int npts,
                         float* __restrict__ fluxes) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  float dn = 0.f, upf = 0.f;
  for (int k = 0; k < ndown; k++) dn = dn + fluxw[k] * rad[(size_t)col_of_ref[k] * ldp + p];
  for (int k = ndown; k < nang; k++) upf = upf + fluxw[k] * rad[(size_t)col_of_ref[k] * ldp + p];
  fluxes[p] = dn;
  fluxes[npts + p] = upf;
}

void launch_sweep_preset(hipStream_t st, const SweepArgs& a, bool up, int col0, int ncol) {
  const int n = (up ? a.nbot : a.ntop) * ncol;
  if (n <= 0) return;
  k_sweep_preset<<<(n + 255) / 256, 256, 0, st>>>(a, up ? 1 : 0, col0, ncol);
}

void launch_sweep_columns(hipStream_t st, const SweepArgs& a, int col0, int ncol) {
  if (ncol <= 0) return;
  k_sweep_columns<<<ncol, SWEEP_THREADS, 0, st>>>(a, col0);
}

void launch_bottom_bc(hipStream_t st, const float* rad, int ldp, const int32_t* bcbot, int nbot,
                      const int32_t* col_of_ref, const float* fluxw, int ndown,
                      const float* dirflux, float albpi, float* bcrad) {
  if (nbot <= 0) return;
  k_bottom_bc<<<(nbot + 63) / 64, 64, 0, st>>>(rad, ldp, bcbot, nbot, col_of_ref, fluxw, ndown,
                                                dirflux, albpi, bcrad);
}

void launch_fluxes(hipStream_t st, const float* rad, int ldp, const int32_t* col_of_ref,
                   const float* fluxw, int ndown, int nang, int npts, float* fluxes) {
  k_fluxes<<<(npts + 63) / 64, 64, 0, st>>>(rad, ldp, col_of_ref, fluxw, ndown, nang, npts, fluxes);
}

}  // namespace at3d
