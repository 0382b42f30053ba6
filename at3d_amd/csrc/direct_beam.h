// Direct solar beam: the slant optical path from a grid point to the top of
// the domain through the property grid (DIRECT_BEAM_PROP shdomsub1.f:4097-4334
// and DIRECT_BEAM_AND_PATHS_PROP shdomsub5.f:1608-2003), shared by the host
// plan (sun paths of the gradient) and the device pass that makes a solve's
// direct flux.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace at3d {

// Trace geometry: the sun direction, the property grid and its delta-M
// scaled total extinction [npx][npy][npz].
struct DirectBeamGeom {
  const float* zl = nullptr;     // [npz] property grid levels
  const float* ext = nullptr;    // [npx * npy * npz] extinction for the direct beam
  int npx = 0, npy = 0, npz = 0, bc = 0, ipdirect = 0;
  int di = 0, dj = 0, dk = 0;
  double cx = 0, cy = 0, cz = 0, cxinv = 0, cyinv = 0, czinv = 0;
  double epss = 0, epsz = 0, xdomain = 0, ydomain = 0, delxd = 0, delyd = 0;
  double solarflux = 0;
  long long max_steps = 100000000;   // the trace gives up (returns false) past this many cells
};

// No per-segment record (the plain direct flux).
struct NoBeamRecord {
  __host__ __device__ void operator()(const int* /*ii*/, const double* /*seg*/) const {}
};

// Slant optical path from (xi, yi, zi) to the domain top in `path`.  `rec`
// receives, per property cell crossed, the eight corner indices and the
// integral of each corner's trilinear weight along the segment (the
// derivative of the path with respect to that corner's extinction).
// Returns false where the reference stops (the ray leaves the grid).
template <class Rec>
__host__ __device__ inline bool direct_beam_trace(const DirectBeamGeom& b, float xi, float yi, float zi,
                                                  double& path, Rec&& rec) {
#pragma clang fp contract(off)
  const int npx = b.npx, npy = b.npy, npz = b.npz, bc = b.bc;
  const float* ZL = b.zl;
  auto bt = [](int v, int bit) { return ((v >> bit) & 1) != 0; };
  double z = zi, x = (double)(xi - 0.0f), y = (double)(yi - 0.0f);
  int il = 0, iu = npz;
  while (iu - il > 1) {
    const int im = (iu + il) / 2;
    if (z >= (double)ZL[im - 1]) il = im; else iu = im;
  }
  int k = il > 1 ? il : 1;
  int i = (int)(x / b.delxd) + 1;
  if (i > npx && fabs(x - b.xdomain) < 0.001 * b.delxd) i = npx;
  if (i < 1 || i > npx) return false;
  int j = (int)(y / b.delyd) + 1;
  if (j > npy && fabs(y - b.ydomain) < 0.001 * b.delyd) j = npy;
  if (j < 1 || j > npy) return false;
  double xe = x, ye = y, ze = z, xp = xe, yp = ye, zp = ze;
  bool constx = bt(b.ipdirect, 0), consty = bt(b.ipdirect, 1);
  if (b.cx == 0.0) constx = true;
  if (b.cy == 0.0) consty = true;
  if (bt(bc, 0) && (fabs(x) < (double)0.01f * b.delxd || fabs(x - (npx - 1) * b.delxd) < (double)0.01f * b.delxd))
    constx = true;
  if (bt(bc, 1) && (fabs(y) < (double)0.01f * b.delyd || fabs(y - (npy - 1) * b.delyd) < (double)0.01f * b.delyd))
    consty = true;
  path = 0.0;
  const float* E = b.ext;
  long long guard = 0;
  while (fabs(ze - (double)ZL[npz - 1]) > b.epsz) {
    if (++guard > b.max_steps) return false;
    int ip1 = i + 1;
    if (i == npx) ip1 = (bt(bc, 0) || bt(bc, 2)) ? npx : 1;
    int jp1 = j + 1;
    if (j == npy) jp1 = (bt(bc, 1) || bt(bc, 3)) ? npy : 1;
    const double x0 = b.delxd * (i - 1), x1 = x0 + b.delxd;
    const double y0 = b.delyd * (j - 1), y1 = y0 + b.delyd;
    if (i < 1 || i > npx || j < 1 || j > npy || k < 1 || k >= npz) return false;
    const double z0 = ZL[k - 1], z1 = ZL[k];
    const int I1 = k + npz * (j - 1) + npz * npy * (i - 1) - 1;
    const int I2 = k + npz * (j - 1) + npz * npy * (ip1 - 1) - 1;
    const int I3 = k + npz * (jp1 - 1) + npz * npy * (i - 1) - 1;
    const int I4 = k + npz * (jp1 - 1) + npz * npy * (ip1 - 1) - 1;
    const double E1 = E[I1], E2 = E[I2], E3 = E[I3], E4 = E[I4];
    const double E5 = E[I1 + 1], E6 = E[I2 + 1], E7 = E[I3 + 1], E8 = E[I4 + 1];
    double sox, soy, soz;
    if (constx) sox = 1.0e30;
    else if (b.cx > 0.0) { sox = (x1 - xe) * b.cxinv; xp = x1; }
    else { sox = (x0 - xe) * b.cxinv; xp = x0; }
    if (consty) soy = 1.0e30;
    else if (b.cy > 0.0) { soy = (y1 - ye) * b.cyinv; yp = y1; }
    else { soy = (y0 - ye) * b.cyinv; yp = y0; }
    if (b.cz > 0.0) { soz = (z1 - ze) * b.czinv; zp = z1; }
    else if (b.cz < 0.0) { soz = (z0 - ze) * b.czinv; zp = z0; }
    else soz = 1.0e30;
    double xoffs = 0.0, yoffs = 0.0, so;
    if (soz <= sox && soz <= soy) {
      so = soz;
      if (!constx) xp = xe + so * b.cx;
      if (!consty) yp = ye + so * b.cy;
      k += b.dk;
    } else if (sox <= soy) {
      so = sox;
      if (!consty) yp = ye + so * b.cy;
      zp = ze + so * b.cz;
      i += b.di;
      if (i == 0) {
        if (bt(bc, 0)) { i = 1; constx = true; }
        else if (bt(bc, 2)) return false;
        else { i = npx; xoffs = b.xdomain; }
      } else if (i >= npx && bt(bc, 2)) {
        return false;
      } else if (i == npx + 1) {
        if (bt(bc, 0)) { i = npx; constx = true; }
        else { i = 1; xoffs = -b.xdomain; }
      }
    } else {
      so = soy;
      if (!constx) xp = xe + so * b.cx;
      zp = ze + so * b.cz;
      j += b.dj;
      if (j == 0) {
        if (bt(bc, 1)) { j = 1; consty = true; }
        else if (bt(bc, 3)) return false;
        else { j = npy; yoffs = b.ydomain; }
      } else if (j >= npy && bt(bc, 3)) {
        return false;
      } else if (j == npy + 1) {
        if (bt(bc, 1)) { j = npy; consty = true; }
        else { j = 1; yoffs = -b.ydomain; }
      }
    }
    if (so < -b.epss) return false;
    so = so > 0.0 ? so : 0.0;
    const double ax = 1.0 / (x1 - x0), ay = 1.0 / (y1 - y0), az = 1.0 / (z1 - z0);
    const double u0 = (xe - x0) * ax, v0 = (ye - y0) * ay, w0 = (ze - z0) * az;
    const double u1 = (xp - x0) * ax, v1 = (yp - y0) * ay, w1 = (zp - z0) * az;
    const double u0m = 1.0 - u0, v0m = 1.0 - v0, w0m = 1.0 - w0;
    const double du = u1 - u0, dv = v1 - v0, dw = w1 - w0;
    double uv = u0 * v0, umv = u0m * v0, uvm = u0 * v0m, umvm = u0m * v0m;
    double uw = u0 * w0, umw = u0m * w0, uwm = u0 * w0m, umwm = u0m * w0m;
    double vw = v0 * w0, vmw = v0m * w0, vwm = v0 * w0m, vmwm = v0m * w0m;
    const double vw0 = vw;
    const double A = (E1 * u0m + E2 * u0) * vmwm + (E3 * u0m + E4 * u0) * vwm +
                     (E5 * u0m + E6 * u0) * vmw + (E7 * u0m + E8 * u0) * vw;
    const double B1 = -du * vmwm - dv * umwm - dw * umvm;
    const double B2 = du * vmwm - dv * uwm - dw * uvm;
    const double B3 = -du * vwm + dv * umwm - dw * umv;
    const double B4 = du * vwm + dv * uwm - dw * uv;
    const double B5 = -du * vmw - dv * umw + dw * umvm;
    const double B6 = du * vmw - dv * uw + dw * uvm;
    const double B7 = -du * vw + dv * umw + dw * umv;
    const double B8 = du * vw + dv * uw + dw * uv;
    const double B = B1 * E1 + B2 * E2 + B3 * E3 + B4 * E4 + B5 * E5 + B6 * E6 + B7 * E7 + B8 * E8;
    vw = dv * dw;
    const double vwu = vw * u0, vwum = vw * u0m;
    uw = du * dw;
    const double uwv = uw * v0, uwvm = uw * v0m;
    uv = du * dv;
    const double uvw = uv * w0, uvwm = uv * w0m;
    const double C1 = +vwum + uwvm + uvwm, C2 = +vwu - uwvm - uvwm, C3 = -vwum + uwv - uvwm;
    const double C4 = -vwu - uwv + uvwm, C5 = -vwum - uwvm + uvw, C6 = -vwu + uwvm - uvw;
    const double C7 = +vwum - uwv - uvw, C8 = +vwu + uwv + uvw;
    const double C = C1 * E1 + C2 * E2 + C3 * E3 + C4 * E4 + C5 * E5 + C6 * E6 + C7 * E7 + C8 * E8;
    const double D = du * dv * dw * (E2 + E3 + E5 + E8 - E1 - E4 - E6 - E7);
    path = path + so * (A + 0.5 * B + 0.3333333333333333 * C + 0.25 * D);
    {
      // integral over t in [0,1] of the corner weight (u0 + t du)(v0 + t dv)(w0 + t dw)
      // (or its complements), times the segment length
      const double g = du * dv * dw;
      const double wa[8] = {u0m * vmwm, u0 * vmwm, u0m * vwm, u0 * vwm,
                            u0m * vmw,  u0 * vmw,  u0m * vw0, u0 * vw0};
      const double wb[8] = {B1, B2, B3, B4, B5, B6, B7, B8};
      const double wc[8] = {C1, C2, C3, C4, C5, C6, C7, C8};
      const double sg[8] = {-1, 1, 1, -1, 1, -1, -1, 1};
      const int ii[8] = {I1, I2, I3, I4, I1 + 1, I2 + 1, I3 + 1, I4 + 1};
      double seg[8];
      for (int c = 0; c < 8; c++)
        seg[c] = so * (wa[c] + 0.5 * wb[c] + 0.3333333333333333 * wc[c] + 0.25 * sg[c] * g);
      rec(ii, seg);
    }
    xe = xp + xoffs;
    ye = yp + yoffs;
    ze = zp;
  }
  return true;
}

// Direct flux (and F0 = direct flux * secmu0) of every grid point on the
// device; err[0] is set to 1 where a trace stops.
void launch_direct_beam(hipStream_t st, const DirectBeamGeom& b, const float* gridpos, int npts, float secmu0,
                        float* dirflux, float* f0, int* err);
// tms [npts][ml1] and ss_coef (CSR by ss_off) times each point's direct flux.
void launch_direct_scale(hipStream_t st, const float* dirflux, int npts, int ml1, float* tms,
                         const int32_t* ss_off, float* ss_coef);

}  // namespace at3d
