// Discrete-ordinate sweeps (BACK_INT_GRID3D_UNPOL, shdomsub1.f:3695-4032).
//
// The reference integrates one ordinate at a time, visiting grid points in its
// sweep order (SWEEPING_ORDER :3261) and tracing each point back to the first
// cell face whose four corner radiances are already known.  The host planner
// replays that order and assigns every point a dependency level per octant
// (all upwind face points of every ordinate of the octant sit on lower
// levels).  On the GPU one launch integrates one level for every ordinate of a
// hemisphere: a thread owns one (point, ordinate) pair, ordinates of a point are
// adjacent lanes, and the DO arrays are stored [point][ordinate] so a point's
// upwind values for all ordinates of an octant share cache lines.  The whole
// level sequence is replayed from a HIP graph.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace at3d {

__device__ inline float sweep_simple(const SweepArgs& a, const OrdGeom& og, int col, int pt) {
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
  const float* se = a.srcext + col;
  const float* rad = a.rad + col;
  const size_t L = a.ldo;
  const double f1 = (1 - u) * (1 - v), f2 = (1 - u) * v, f3 = u * (1 - v), f4 = u * v;
  const double ext0 = f1 * ext[i1] + f2 * ext[i2] + f3 * ext[i3] + f4 * ext[i4];
  const double srcext0 = f1 * se[i1 * L] + f2 * se[i2 * L] + f3 * se[i3 * L] + f4 * se[i4 * L];
  double src, absc, trc;
  cell_integral(ext0, (double)ext[pt], srcext0, (double)se[pt * L], so, src, absc, trc);
  double radv = 0.0 + 1.0 * src * absc;
  const double transmit = 1.0 * trc;
  const double rad0 = f1 * rad[i1 * L] + f2 * rad[i2 * L] + f3 * rad[i3 * L] + f4 * rad[i4 * L];
  radv = radv + transmit * rad0;
  return (float)radv;
}

// General characteristic: trace through cells until a face with known radiances.
__device__ inline float sweep_general(const SweepArgs& a, const OrdGeom& og, int col, int o, int pt,
                                      int icell, int L) {
  const int presetbit = og.bitz ? 2 : 1;
  const int32_t* rank = a.rank + (size_t)o * a.npts;
  const int32_t* level = a.level + (size_t)o * a.npts;
  const int rp = rank[pt];
  const float* ext = a.ext;
  const float* se = a.srcext + col;
  const float* rad = a.rad + col;
  const size_t LD = a.ldo;
  double xe = a.g.gridpos[pt * 3], ye = a.g.gridpos[pt * 3 + 1], ze = a.g.gridpos[pt * 3 + 2];
  double ext1 = ext[pt], srcext1 = se[pt * LD];
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
    const double srcext0 = f1 * se[t.i[0] * LD] + f2 * se[t.i[1] * LD] + f3 * se[t.i[2] * LD] +
                           f4 * se[t.i[3] * LD];
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
      const double rad0 = f1 * rad[t.i[0] * LD] + f2 * rad[t.i[1] * LD] + f3 * rad[t.i[2] * LD] +
                          f4 * rad[t.i[3] * LD];
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

__global__ __launch_bounds__(256) void k_sweep_level(SweepArgs a, LevelDesc d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.item_off[d.nslot]) return;
  int s = 0;
  while (s + 1 < d.nslot && i >= d.item_off[s + 1]) s++;
  const int rel = i - d.item_off[s];
  const int j = rel / d.nord[s];
  const int r = rel - j * d.nord[s];
  const int col = d.col_base[s] + r;
  const int code = a.lvl_pts[d.pts_off[s] + j];
  const OrdGeom og = a.ordg[col];
  float v;
  int pt;
  if (code >= 0 && !a.force_general) {
    pt = code;
    v = sweep_simple(a, og, col, pt);
  } else {
    pt = code >= 0 ? code : ~code;
    v = sweep_general(a, og, col, d.oct[s], pt, a.lvl_cell[d.pts_off[s] + j], d.L);
  }
  a.rad[(size_t)pt * a.ldo + col] = v;
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
    a.rad[(size_t)a.bcbot[b] * a.ldo + c] = a.bcrad_bot[b];
  else
    a.rad[(size_t)a.bctop[b] * a.ldo + c] = a.skyrad[c];
}

// FIXED_LAMBERTIAN_BOUNDARY (shdomsub1.f:2438-2480) with the downwelling
// hemispheric flux summed in the reference's ordinate order (:2118-2135).
__global__ void k_bottom_bc(const float* __restrict__ rad, int ldo, const int32_t* __restrict__ bcbot,
                            int nbot, const int32_t* __restrict__ col_of_ref,
                            const float* __restrict__ fluxw, int ndown,
                            const float* __restrict__ dirflux, float albpi,
                            float* __restrict__ bcrad) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbot) return;
  const int p = bcbot[b];
  const float* row = rad + (size_t)p * ldo;
  float f = 0.f;
  for (int k = 0; k < ndown; k++) f = f + fluxw[k] * row[col_of_ref[k]];
  bcrad[b] = albpi * (dirflux[p] + f);
}

__global__ void k_fluxes(const float* __restrict__ rad, int ldo, const int32_t* __restrict__ col_of_ref,
                         const float* __restrict__ fluxw, int ndown, int nang, int npts,
                         float* __restrict__ fluxes) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  const float* row = rad + (size_t)p * ldo;
  float dn = 0.f, upf = 0.f;
  for (int k = 0; k < ndown; k++) dn = dn + fluxw[k] * row[col_of_ref[k]];
  for (int k = ndown; k < nang; k++) upf = upf + fluxw[k] * row[col_of_ref[k]];
  fluxes[p] = dn;
  fluxes[npts + p] = upf;
}

void launch_sweep_preset(hipStream_t st, const SweepArgs& a, bool up, int col0, int ncol) {
  const int n = (up ? a.nbot : a.ntop) * ncol;
  if (n <= 0) return;
  k_sweep_preset<<<(n + 255) / 256, 256, 0, st>>>(a, up ? 1 : 0, col0, ncol);
}

void launch_sweep_level(hipStream_t st, const SweepArgs& a, const LevelDesc& d) {
  const int n = d.item_off[d.nslot];
  if (n <= 0) return;
  k_sweep_level<<<(n + 255) / 256, 256, 0, st>>>(a, d);
}

void launch_bottom_bc(hipStream_t st, const float* rad, int ldo, const int32_t* bcbot, int nbot,
                      const int32_t* col_of_ref, const float* fluxw, int ndown,
                      const float* dirflux, float albpi, float* bcrad) {
  if (nbot <= 0) return;
  k_bottom_bc<<<(nbot + 63) / 64, 64, 0, st>>>(rad, ldo, bcbot, nbot, col_of_ref, fluxw, ndown,
                                                dirflux, albpi, bcrad);
}

void launch_fluxes(hipStream_t st, const float* rad, int ldo, const int32_t* col_of_ref,
                   const float* fluxw, int ndown, int nang, int npts, float* fluxes) {
  k_fluxes<<<(npts + 63) / 64, 64, 0, st>>>(rad, ldo, col_of_ref, fluxw, ndown, nang, npts, fluxes);
}

}  // namespace at3d
