// Radiance rendering: one ray per lane (RENDER shdomsub4.f:93-286 with
// INTEGRATE_1RAY shdomsub2.f:2311-2743).
//
// The direction-dependent source at the grid points,
// COMPUTE_SOURCE_1CELL_UNPOL (shdomsub2.f:3046-3192) =
//   ext * ( sum_J S_J Y_J(dir)                       [spherical-harmonic source]
//           - sum_{l<=LS} T_l sum_{J in l} Y_J(dir) Ysun_J   [truncated single scatter]
//           + sum_pairs c * P_table(Theta) )          [exact single scatter, TMS]
// is computed once per distinct ray direction for every grid point (an MFMA
// product [points x nlm] x [nlm x directions] plus a per-point epilogue), so
// the ray kernel only gathers cell corners and integrates.
#include <hip/hip_runtime.h>
#include <math.h>

#include "kernels.h"
#include "ray_common.h"

namespace at3d {

__device__ inline int nterms_l(int l, int mm) {
  return (l <= mm) ? l * (l + 1) + l + 1 : (2 * mm + 1) * l - mm * mm + mm + 1;
}

// dsrc[p][d] = ext[p] * (raw[p][d] - sum_{l<=LS(p)} T[p][l] G[d][l] + sum_pairs c * ss[d][iph])
__global__ void k_dir_source_finish(const float* raw, int ldd, int ndir,
                                    const float* __restrict__ ext, const int* __restrict__ ns,
                                    const int* __restrict__ lofj, const float* __restrict__ tms,
                                    const float* __restrict__ G, const int32_t* __restrict__ ss_off,
                                    const int32_t* __restrict__ ss_iph,
                                    const float* __restrict__ ss_coef, const float* __restrict__ ssd,
                                    int numphase, int ml, int npts, int single_scatter, int tms_on,
                                    float* dsrc) {
  // raw and dsrc may be the same buffer: each element is read and then
  // written by one thread.
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts * ndir) return;
  const int p = i / ndir, d = i - p * ndir;
  float v = single_scatter ? 0.0f : raw[(size_t)p * ldd + d];
  if (tms_on) {
    const int nsp = ns[p];
    if (!single_scatter && nsp > 0) {
      const int ls = lofj[nsp - 1];
      for (int l = 0; l <= ls; l++) v = v - tms[(size_t)p * (ml + 1) + l] * G[d * (ml + 1) + l];
    }
    for (int q = ss_off[p]; q < ss_off[p + 1]; q++) v = v + ss_coef[q] * ssd[d * numphase + ss_iph[q]];
  }
  dsrc[(size_t)p * ldd + d] = v * ext[p];
}



// The polarized direction source (COMPUTE_SOURCE_1CELL, shdomsub2.f:2868-3043)
// for I, Q and U: each block's spherical-harmonic sum less its truncated
// single scatter (a1 for I, b1 for Q and U, with their direction factors) plus
// the exact single scatter of the tabulated I, rotated Q and U.
__global__ void k_dir_source_finish_pol(const float* raw, int ldd, int cst, int ndir,
                                        const float* __restrict__ ext, const int* __restrict__ ns,
                                        const int* __restrict__ lofj, const float* __restrict__ tms,
                                        const float* __restrict__ tms5, const float* __restrict__ G3,
                                        const int32_t* __restrict__ ss_off, const int32_t* __restrict__ ss_iph,
                                        const float* __restrict__ ss_coef, const float* __restrict__ ssd,
                                        const float* __restrict__ ssq, const float* __restrict__ ssu,
                                        int numphase, int ml, int npts, int single_scatter, int tms_on,
                                        float* dsrc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts * ndir) return;
  const int p = i / ndir, d = i - p * ndir;
  const size_t row = (size_t)p * ldd;
  float v[3];
  for (int k = 0; k < 3; k++) v[k] = single_scatter ? 0.0f : raw[row + k * cst + d];
  if (tms_on) {
    const int nsp = ns[p];
    if (!single_scatter && nsp > 0) {
      const int ls = lofj[nsp - 1];
      const size_t gd = (size_t)ndir * (ml + 1);
      for (int l = 0; l <= ls; l++) {
        const float t1 = tms[(size_t)p * (ml + 1) + l], t5 = tms5[(size_t)p * (ml + 1) + l];
        v[0] = v[0] - t1 * G3[d * (ml + 1) + l];
        v[1] = v[1] - t5 * G3[gd + d * (ml + 1) + l];
        v[2] = v[2] - t5 * G3[2 * gd + d * (ml + 1) + l];
      }
    }
    for (int q = ss_off[p]; q < ss_off[p + 1]; q++) {
      const float c = ss_coef[q];
      const int t = d * numphase + ss_iph[q];
      v[0] = v[0] + c * ssd[t];
      v[1] = v[1] + c * ssq[t];
      v[2] = v[2] + c * ssu[t];
    }
  }
  for (int k = 0; k < 3; k++) dsrc[row + k * cst + d] = v[k] * ext[p];
}

__device__ inline float trilin(const float* v8, double u, double v, double w) {
#pragma clang fp contract(off)
  return (float)((1 - w) * ((1 - v) * ((1 - u) * v8[0] + u * v8[1]) + v * ((1 - u) * v8[2] + u * v8[3])) +
                 w * ((1 - v) * ((1 - u) * v8[4] + u * v8[5]) + v * ((1 - u) * v8[6] + u * v8[7])));
}

// NST = 3: I, Q and U integrated along the ray with the same transmission;
// only I's source is clamped at zero, and Q and U have zero boundary
// radiances (the sky is unpolarized, the Lambertian surface depolarizes).
template <int NST>
__global__ __launch_bounds__(128) void k_render(RayArgs a) {
#pragma clang fp contract(off)
  const int ir = blockIdx.x * blockDim.x + threadIdx.x;
  if (ir >= a.nrays) return;
  const double PI = acos(-1.0);
  const float* ZG = a.zgrid;
  const float* P = a.g.gridpos;
  double x0 = a.camx[ir], y0 = a.camy[ir], z0 = a.camz[ir];
  const double mu2 = a.cammu[ir], phi2 = a.camphi[ir];
  const double muray = -mu2, phiray = phi2 - PI;
  // extrapolate a sensor above the domain to the top (RENDER :223-236)
  if (z0 > (double)ZG[a.nz - 1]) {
    if (muray >= 0.0) {
      for (int k = 0; k < NST; k++) a.out[(size_t)ir * NST + k] = 0.0f;
      return;
    }
    const double r = ((double)ZG[a.nz - 1] - z0) / muray;
    x0 = x0 + r * sqrt(1 - muray * muray) * cos(phiray);
    y0 = y0 + r * sqrt(1 - muray * muray) * sin(phiray);
    z0 = ZG[a.nz - 1];
  }
  const int d = a.dir_of_ray[ir];
  const float topv = a.topval[ir];
  const float* dsrc = a.dsrc + d;
  const size_t LD = a.ldd;
  const double eps = (double)(1.0e-5f * (P[a.g.gridptr[7] * 3 + 2] - P[a.g.gridptr[0] * 3 + 2]));
  const int maxcross = 500 * max(a.nx, max(a.ny, a.nz));
  double cx = sqrt(1.0 - mu2 * mu2) * cos(phi2 - PI);
  double cy = sqrt(1.0 - mu2 * mu2) * sin(phi2 - PI);
  double cz = -mu2, cxinv, cyinv, czinv;
  if (fabs(cx) > (double)1.0e-6f) cxinv = 1.0 / cx; else { cx = 0.0; cxinv = 1.0e6; }
  if (fabs(cy) > (double)1.0e-6f) cyinv = 1.0 / cy; else { cy = 0.0; cyinv = 1.0e6; }
  if (fabs(cz) > (double)1.0e-6f) czinv = 1.0 / cz; else { cz = 0.0; czinv = 1.0e6; }
  const int bitx = cx < 0.0, bity = cy < 0.0, bitz = cz < 0.0;
  const int ioct = bitx + 2 * bity + 4 * bitz;
  const float xm = 0.5f * (a.xgrid[0] + a.xgrid[a.nx - 1]);
  const float ym = 0.5f * (a.ygrid[0] + a.ygrid[a.ny - 1]);
  double xe = x0, ye = y0, ze = z0;
  int icell = locate_cell_ray(a, xe, ye, ze);   // (wraps xe, ye into a periodic domain)
  if (icell >= 0 && icell < a.g.ncells) icell = locate_in_tree(a.g, icell, xe, ye, ze);
  if (icell < 0 || icell >= a.g.ncells) {
    atomicOr(a.errflag, 4);
    for (int k = 0; k < NST; k++) a.out[(size_t)ir * NST + k] = 0.0f;
    return;
  }
  double radiance[NST], transmit = 1.0;
  for (int k = 0; k < NST; k++) radiance[k] = 0.0;
  int ngrid = 0;
  bool validrad = false;
  float e8[8], s8[NST][8];
  float srcext1[NST], ext1 = 0.f;
  while (!validrad && icell >= 0) {
    ngrid++;
    const int* gp = a.g.gridptr + (size_t)icell * 8;
    for (int c = 0; c < 8; c++) {
      const int q = gp[c];
      e8[c] = a.ext[q];
      for (int k = 0; k < NST; k++) s8[k][c] = dsrc[(size_t)q * LD + k * a.cst];
    }
    const int ipt1 = gp[0], ipt2 = gp[7];
    const double delx = (double)(P[ipt2 * 3] - P[ipt1 * 3]);
    const double dely = (double)(P[ipt2 * 3 + 1] - P[ipt1 * 3 + 1]);
    const double delz = (double)(P[ipt2 * 3 + 2] - P[ipt1 * 3 + 2]);
    const double invdelx = delx <= 0.0 ? 1.0 : 1.0 / delx;
    const double invdely = dely <= 0.0 ? 1.0 : 1.0 / dely;
    const double invdelz = 1.0 / delz;
    const double px1 = P[ipt1 * 3], py1 = P[ipt1 * 3 + 1], pz1 = P[ipt1 * 3 + 2];
    double u = (xe - px1) * invdelx, v = (ye - py1) * invdely, w = (ze - pz1) * invdelz;
    // the entry values: interpolated in this cell (CORRECTINTERPOLATE, and
    // always in the first), else those the ray left the last cell with
    if (a.correct_interp || ngrid == 1) {
      for (int k = 0; k < NST; k++) srcext1[k] = trilin(s8[k], u, v, w);
      srcext1[0] = fmaxf(0.0f, srcext1[0]);
      ext1 = trilin(e8, u, v, w);
    }
    const int8_t fl = a.g.cellflags[icell];
    const bool ipinx = (fl & 1) && !(a.open_x && ((cx > 0 && xe < xm) || (cx < 0 && xe > xm)));
    const bool ipiny = (fl & 2) && !(a.open_y && ((cy > 0 && ye < ym) || (cy < 0 && ye > ym)));
    const int iopp = gp[7 - ioct];
    const double sox = ipinx ? 1.0e20 : ((double)P[iopp * 3] - xe) * cxinv;
    const double soy = ipiny ? 1.0e20 : ((double)P[iopp * 3 + 1] - ye) * cyinv;
    const double soz = ((double)P[iopp * 3 + 2] - ze) * czinv;
    double so = fmin(sox, fmin(soy, soz));
    if (so < -eps) {
      atomicOr(a.errflag, 1);
      break;
    }
    double xn = xe + so * cx, yn = ye + so * cy, zn = ze + so * cz;
    u = (xn - px1) * invdelx;
    v = (yn - py1) * invdely;
    w = (zn - pz1) * invdelz;
    const float extn = trilin(e8, u, v, w);
    const double taugrid = so * 0.5 * (double)(ext1 + extn);
    const int ntau = max(1, 1 + (int)(taugrid / a.tautol));
    const double dels = so / ntau;
    for (int it = 1; it <= ntau; it++) {
      const double s = it * dels;
      u = (xe + s * cx - px1) * invdelx;
      v = (ye + s * cy - py1) * invdely;
      w = (ze + s * cz - pz1) * invdelz;
      const float ext0 = trilin(e8, u, v, w);
      float srcext0[NST];
      for (int k = 0; k < NST; k++) srcext0[k] = trilin(s8[k], u, v, w);
      srcext0[0] = fmaxf(0.0f, srcext0[0]);
      const double ext = (double)(0.5f * (ext0 + ext1));
      double absc, trc;
      if (ext != 0.0) {
        const double tau = ext * dels;
        absc = tau * (1.0 - 0.5 * tau * (1.0 - (double)0.33333333333f * tau));
        trc = 1.0 - absc;
        for (int k = 0; k < NST; k++) {
          const double src = ((double)(0.5f * (srcext0[k] + srcext1[k])) +
                              (double)(0.08333333333f * (ext0 * srcext1[k] - ext1 * srcext0[k])) * dels *
                                  (1.0 - (double)(0.05f * (ext1 - ext0)) * dels)) / ext;
          radiance[k] = radiance[k] + transmit * src * absc;
        }
      } else {
        absc = 0.0;
        trc = 1.0;
      }
      transmit = transmit * trc;
      ext1 = ext0;
      for (int k = 0; k < NST; k++) srcext1[k] = srcext0[k];
    }
    int iface, jface;
    bool openbcface;
    if (sox <= soz && sox <= soy) {
      iface = 1 - bitx;
      jface = 1;
      openbcface = (fl & 1) && a.open_x;
    } else if (soy <= soz) {
      iface = 3 - bity;
      jface = 2;
      openbcface = (fl & 2) && a.open_y;
    } else {
      iface = 5 - bitz;
      jface = 3;
      openbcface = false;
    }
    // a face onto smaller cells continues down the neighbour's tree, and the
    // boundary interpolation takes that smaller face (NEXT_CELL)
    const int nbv = a.g.neighptr[(size_t)icell * 6 + iface];
    const int inext = nb_smaller(nbv) ? next_cell_down(a.g, nb_cell(nbv), iface, jface, xn, yn, zn) : nbv;
    int kface, ic;
    bool zface = iface >= 4;
    if (!openbcface && !nb_smaller(nbv)) {
      kface = iface;
      ic = icell;
    } else {
      kface = iface ^ 1;
      ic = inext;
      zface = false;
    }
    if (inext >= 0) {
      const int q = a.g.gridptr[(size_t)inext * 8 + ioct];
      if (jface == 1) xn = P[q * 3];
      else if (jface == 2) yn = P[q * 3 + 1];
      else zn = P[q * 3 + 2];
    }
    if (transmit < a.transcut || ngrid > maxcross) {
      validrad = true;
    } else if (inext < 0 && zface) {
      validrad = true;
      // FIND_BOUNDARY_RADIANCE (shdomsub2.f:2748-2863), Lambertian surface
      const int8_t tab[6][4] = {{0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5},
                                {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};
      float X[4], Y[4], R[4];
      for (int j = 0; j < 4; j++) {
        const int ip = a.g.gridptr[(size_t)ic * 8 + tab[kface][j]];
        X[j] = P[ip * 3];
        Y[j] = P[ip * 3 + 1];
        if ((float)mu2 < 0.0f) {
          R[j] = topv;
        } else {
          const int b = a.bot_index[ip];
          if (b < 0) atomicOr(a.errflag, 2);
          R[j] = b < 0 ? 0.0f : a.bcrad_bot[b];
        }
      }
      const float ub = (X[1] - X[0] > 0.0f) ? (float)((xn - X[0]) / (double)(X[1] - X[0])) : 0.0f;
      const float vb = (Y[2] - Y[0] > 0.0f) ? (float)((yn - Y[0]) / (double)(Y[2] - Y[0])) : 0.0f;
      const float radbnd = (1 - ub) * (1 - vb) * R[0] + ub * (1 - vb) * R[1] +
                           (1 - ub) * vb * R[2] + ub * vb * R[3];
      if (!a.nosurface) radiance[0] = radiance[0] + transmit * (double)radbnd;
    } else {
      icell = inext;
    }
    xe = xn;
    ye = yn;
    ze = zn;
  }
  for (int k = 0; k < NST; k++) a.out[(size_t)ir * NST + k] = (float)radiance[k];
}

// Surface radiance from the final fluxes (RENDER :201-209).
__global__ void k_render_bottom(const float* __restrict__ fluxes, const float* __restrict__ dirflux,
                                const int32_t* __restrict__ bcbot, int nbot, float albpi,
                                float* __restrict__ bcrad) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbot) return;
  const int p = bcbot[b];
  bcrad[b] = albpi * (dirflux[p] + fluxes[p]);
}

void launch_dir_source_finish(hipStream_t st, const float* raw, int ldd, int ndir, const float* ext,
                              const int* ns, const int* lofj, const float* tms, const float* G,
                              const int32_t* ss_off, const int32_t* ss_iph, const float* ss_coef,
                              const float* ssd, int numphase, int ml, int npts, int single_scatter,
                              int tms_on, float* dsrc) {
  const int n = npts * ndir;
  k_dir_source_finish<<<(n + 255) / 256, 256, 0, st>>>(raw, ldd, ndir, ext, ns, lofj, tms, G, ss_off,
                                                        ss_iph, ss_coef, ssd, numphase, ml, npts,
                                                        single_scatter, tms_on, dsrc);
}

void launch_render(hipStream_t st, const RayArgs& a) {
  if (a.nst == 3)
    k_render<3><<<(a.nrays + 127) / 128, 128, 0, st>>>(a);
  else
    k_render<1><<<(a.nrays + 127) / 128, 128, 0, st>>>(a);
}

void launch_dir_source_finish_pol(hipStream_t st, const float* raw, int ldd, int cst, int ndir, const float* ext,
                                  const int* ns, const int* lofj, const float* tms, const float* tms5,
                                  const float* G3, const int32_t* ss_off, const int32_t* ss_iph,
                                  const float* ss_coef, const float* ssd, const float* ssq, const float* ssu,
                                  int numphase, int ml, int npts, int single_scatter, int tms_on, float* dsrc) {
  const int n = npts * ndir;
  k_dir_source_finish_pol<<<(n + 255) / 256, 256, 0, st>>>(raw, ldd, cst, ndir, ext, ns, lofj, tms, tms5, G3,
                                                            ss_off, ss_iph, ss_coef, ssd, ssq, ssu, numphase,
                                                            ml, npts, single_scatter, tms_on, dsrc);
}

void launch_render_bottom(hipStream_t st, const float* fluxes, const float* dirflux,
                          const int32_t* bcbot, int nbot, float albpi, float* bcrad) {
  if (nbot <= 0) return;
  k_render_bottom<<<(nbot + 255) / 256, 256, 0, st>>>(fluxes, dirflux, bcbot, nbot, albpi, bcrad);
}

}  // namespace at3d
