// Levis approximate gradient on the GPU (LEVISAPPROX_GRADIENT adjoint branch,
// shdomsub4.f:288-809).
//
// The reference walks every ray backwards through the grid, building per
// cell corner the "gradient source" GRAD8 (COMPUTE_SOURCE_GRAD_1CELL
// :1546-2042), and adds transmission x absorption x GRAD8 into the property
// grid gradient; it stores every sub-step of the ray to add the radiance term
// afterwards (COMPUTE_RADIANCE_DERIVATIVE_ADJOINT :4037-4114).  Here:
//
//  * GRAD8 is constant within a cell, so the ray kernel only accumulates, per
//    (grid point, ray direction), the scalar weight sum T ABS alpha(corner)
//    (alpha = the second-order quadrature factor of that corner) and, per grid
//    point, the radiance-term weight and the direct-beam weight; GRAD8 is then
//    contracted once per (point, direction) by k_grad_points.
//  * The ray is traced twice: the first pass gives the total radiance, the
//    second recomputes each sub-step and gets the future radiance of the
//    reference's backward cumulative sum as total minus the running prefix,
//    so no per-ray storage is needed.
#include <hip/hip_runtime.h>
#include <math.h>

#include "kernels.h"
#include "ray_common.h"

namespace at3d {

// Per-direction tables for the gradient: dsrc = source x extinction used along
// the ray (single scatter only if requested), srcpre = the source before the
// extinction factor, ss8 = exact single scatter x extinction.
__global__ void k_dir_source_grad(const float* raw, int ldd, int ndir, const float* __restrict__ ext,
                                  const int* __restrict__ ns, const int* __restrict__ lofj,
                                  const float* __restrict__ tms, const float* __restrict__ G,
                                  const int32_t* __restrict__ ss_off, const int32_t* __restrict__ ss_iph,
                                  const float* __restrict__ ss_coef, const float* __restrict__ ssd,
                                  int numphase, int ml, int npts, int single_scatter,
                                  float* __restrict__ dsrc, float* __restrict__ srcpre,
                                  float* __restrict__ ss8) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts * ndir) return;
  const int p = i / ndir, d = i - p * ndir;
  // same operation order as the render's direction source, so the forward
  // radiances of a gradient call equal at3d_render's
  float v = raw[(size_t)p * ldd + d];
  const int nsp = ns[p];
  if (nsp > 0) {
    const int ls = lofj[nsp - 1];
    for (int l = 0; l <= ls; l++) v = v - tms[(size_t)p * (ml + 1) + l] * G[d * (ml + 1) + l];
  }
  float ss = 0.0f;
  for (int q = ss_off[p]; q < ss_off[p + 1]; q++) {
    const float t = ss_coef[q] * ssd[d * numphase + ss_iph[q]];
    v = v + t;
    ss = ss + t;
  }
  const float e = ext[p];
  const size_t o = (size_t)p * ldd + d;
  srcpre[o] = v;
  ss8[o] = ss * e;
  dsrc[o] = single_scatter ? ss * e : v * e;
}

// GET_INTERP_KERNEL (shdomsub5.f:1497-1537)
__device__ inline void interp_kernel(const GridTopo& g, int icell, double x, double y, double z,
                                     double* F) {
#pragma clang fp contract(off)
  const int* gp = g.gridptr + (size_t)icell * 8;
  const float* P = g.gridpos;
  const int i1 = gp[0], i2 = gp[7];
  const double delx = (double)(P[i2 * 3] - P[i1 * 3]);
  const double dely = (double)(P[i2 * 3 + 1] - P[i1 * 3 + 1]);
  const double delz = (double)(P[i2 * 3 + 2] - P[i1 * 3 + 2]);
  const double ix = delx <= 0.0 ? 1.0 : 1.0 / delx;
  const double iy = dely <= 0.0 ? 1.0 : 1.0 / dely;
  const double iz = 1.0 / delz;
  const double u = (x - (double)P[i1 * 3]) * ix;
  const double v = (y - (double)P[i1 * 3 + 1]) * iy;
  const double w = (z - (double)P[i1 * 3 + 2]) * iz;
  F[0] = (1 - w) * (1 - v) * (1 - u);
  F[1] = (1 - w) * (1 - v) * u;
  F[2] = (1 - w) * v * (1 - u);
  F[3] = (1 - w) * v * u;
  F[4] = w * (1 - v) * (1 - u);
  F[5] = w * (1 - v) * u;
  F[6] = w * v * (1 - u);
  F[7] = w * v * u;
}

__device__ inline float dot8(const double* F, const float* v) {
#pragma clang fp contract(off)
  double s = F[0] * v[0];
  for (int k = 1; k < 8; k++) s = s + F[k] * v[k];
  return (float)s;
}

__device__ inline void atomic_add_f64(double* p, double v) {
  if (v != 0.0) atomicAdd(p, v);
}

// One ray: pass 0 returns the total radiance; pass 1 accumulates the adjoint
// weights.  Returns false on a reference error (errflag set).
__device__ bool adjoint_ray(const AdjArgs& A, int ir, int pass, double rtot, double& radout) {
#pragma clang fp contract(off)
  const RayArgs& a = A.r;
  const double PI = acos(-1.0);
  const float* ZG = a.zgrid;
  const float* P = a.g.gridpos;
  double x0 = a.camx[ir], y0 = a.camy[ir], z0 = a.camz[ir];
  const double mu2 = a.cammu[ir], phi2 = a.camphi[ir];
  const double muray = -mu2, phiray = phi2 - PI;
  radout = 0.0;
  if (z0 > (double)ZG[a.nz - 1]) {
    if (muray >= 0.0) return true;
    const double r = ((double)ZG[a.nz - 1] - z0) / muray;
    x0 = x0 + r * sqrt(1 - muray * muray) * cos(phiray);
    y0 = y0 + r * sqrt(1 - muray * muray) * sin(phiray);
    z0 = ZG[a.nz - 1];
  }
  const int d = a.dir_of_ray[ir];
  const double wgt = A.adj_weight[ir];
  const bool jac = A.jac_weight != nullptr;
  const double jwt = jac ? A.jac_weight[ir] : 0.0;
  double* jsrc = jac ? A.jsrc + (size_t)ir * A.nrel_s : nullptr;
  double* jrad = jac ? A.jrad + (size_t)ir * A.nrel_s : nullptr;
  double* jbeam = jac ? A.jbeam + (size_t)ir * A.nrel_b : nullptr;
  const float topv = a.topval[ir];
  const size_t LD = a.ldd;
  const double eps = (double)(1.0e-5f * (P[a.g.gridptr[7] * 3 + 2] - P[a.g.gridptr[0] * 3 + 2]));
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
    return false;
  }
  const bool ess = A.exact_single_scatter != 0;
  double transmit = 1.0;
  double prefix = 0.0;       // sum of F_n over recorded intervals so far
  int npassed = 0;
  bool validrad = false;
  // the last recorded interval is finished only when the next one (or the
  // end of the ray) gives its future radiance
  bool have_prev = false;
  double pv_dels = 0, pv_t = 0, pv_abs = 0, pv_f = 0;
  double pv_i0[8], pv_i1[8];
  int pv_pts[8];
  double rend = 0.0;   // future radiance after the last interval (sentinel)
  auto finish_prev = [&](double future_next) {
    // COMPUTE_RADIANCE_DERIVATIVE_ADJOINT for the previous interval
    double ext0 = 0.0, ext1 = 0.0;
    for (int k = 0; k < 8; k++) {
      ext0 = ext0 + (double)a.ext[pv_pts[k]] * pv_i0[k];
      ext1 = ext1 + (double)a.ext[pv_pts[k]] * pv_i1[k];
    }
    const double ext = (double)(0.5f) * (ext0 + ext1);
    if (ext == 0.0) return;
    const double rad_k = (rtot - (prefix - pv_f)) / pv_t;   // PASSEDRAD(k)
    const double rad_k1 = future_next;                      // PASSEDRAD(k+1)
    auto term = [&](double w, int k) {
      const double g0 = -w * rad_k1 * pv_i0[k];
      const double g1 = -w * rad_k * pv_i1[k];
      const double ag = ((double)0.5f * (g0 + g1) +
                         (double)0.08333333333f * (ext0 * g1 - ext1 * g0) * pv_dels *
                             (1.0 - (double)0.05f * (ext1 - ext0) * pv_dels)) / ext;
      return ag * pv_t * pv_abs;
    };
    for (int k = 0; k < 8; k++) {
      atomic_add_f64(&A.erad[pv_pts[k]], term(wgt, k));
      if (jac) {
        const int r = A.rel_s[pv_pts[k]];
        if (r >= 0) jrad[r] += term(jwt, k);
      }
    }
  };
  float e8[8], s8[8], ss8[8];
  // the reference sets no limit on the cells a ray crosses; a ray that would
  // never leave the grid (horizontal through a periodic, transparent layer)
  // is stopped with an error instead of running forever
  const int maxcells = 1000 * max(a.nx, max(a.ny, a.nz));
  int ngrid = 0;
  while (!validrad && icell >= 0) {
    if (++ngrid > maxcells) {
      atomicOr(a.errflag, 32);
      return false;
    }
    const int* gp = a.g.gridptr + (size_t)icell * 8;
    int pts[8];
    for (int c = 0; c < 8; c++) {
      const int q = gp[c];
      pts[c] = q;
      e8[c] = a.ext[q];
      s8[c] = a.dsrc[(size_t)q * LD + d];
      ss8[c] = A.ss8[(size_t)q * LD + d];
    }
    double fc[8];
    interp_kernel(a.g, icell, xe, ye, ze, fc);
    float srcext1 = fmaxf(0.0f, dot8(fc, s8));
    float ext1 = dot8(fc, e8);
    float sg1[8];
    for (int n = 0; n < 8; n++) sg1[n] = fmaxf(0.0f, (float)(fc[n] * (double)ss8[n]));
    double fc1[8];
    for (int n = 0; n < 8; n++) fc1[n] = fc[n];
    const int8_t fl = a.g.cellflags[icell];
    const bool ipinx = (fl & 1) && !(a.open_x && ((cx > 0 && xe < xm) || (cx < 0 && xe > xm)));
    const bool ipiny = (fl & 2) && !(a.open_y && ((cy > 0 && ye < ym) || (cy < 0 && ye > ym)));
    const int iopp = gp[7 - ioct];
    const double sox = ipinx ? 1.0e20 : ((double)P[iopp * 3] - xe) * cxinv;
    const double soy = ipiny ? 1.0e20 : ((double)P[iopp * 3 + 1] - ye) * cyinv;
    const double soz = ((double)P[iopp * 3 + 2] - ze) * czinv;
    const double so = fmin(sox, fmin(soy, soz));
    if (so < -eps) {
      atomicOr(a.errflag, 1);
      return false;
    }
    double xn = xe + so * cx, yn = ye + so * cy, zn = ze + so * cz;
    double fcn[8];
    interp_kernel(a.g, icell, xn, yn, zn, fcn);
    const float extn = dot8(fcn, e8);
    const double taugrid = so * 0.5 * (double)(ext1 + extn);
    const int ntau = max(1, 1 + (int)(taugrid / a.tautol));
    const double dels = so / ntau;
    double beta[8], sss[8];
    for (int n = 0; n < 8; n++) { beta[n] = 0.0; sss[n] = 0.0; }
    for (int it = 1; it <= ntau; it++) {
      const double s = it * dels;
      interp_kernel(a.g, icell, xe + s * cx, ye + s * cy, ze + s * cz, fc);
      float srcext0 = dot8(fc, s8);
      const float ext0 = (it != ntau) ? dot8(fc, e8) : extn;
      srcext0 = fmaxf(0.0f, srcext0);
      float sg0[8];
      for (int n = 0; n < 8; n++) sg0[n] = fmaxf(0.0f, (float)(fc[n] * (double)ss8[n]));
      const double ext = (double)(0.5f * (ext0 + ext1));
      double trc = 1.0;
      if (ext != 0.0) {
        const double tau = ext * dels;
        const double absc = tau * (1.0 - 0.5 * tau * (1.0 - (double)0.33333333333f * tau));
        trc = 1.0 - absc;
        const double q2 = dels * (1.0 - (double)(0.05f * (ext1 - ext0)) * dels);
        const double src = ((double)(0.5f * (srcext0 + srcext1)) +
                            (double)(0.08333333333f * (ext0 * srcext1 - ext1 * srcext0)) * q2) / ext;
        if (pass == 1) {
          for (int n = 0; n < 8; n++) {
            const float f0 = (float)fc[n], f1 = (float)fc1[n];
            const double al = ((double)(0.5f * (f0 + f1)) +
                               (double)(0.08333333333f * (ext0 * f1 - ext1 * f0)) * q2) / ext;
            beta[n] = beta[n] + transmit * al * absc;
          }
          if (ess) {
            for (int n = 0; n < 8; n++)
              sss[n] = sss[n] + transmit * absc *
                                    ((double)(0.5f * (sg0[n] + sg1[n])) +
                                     (double)(0.08333333333f * (ext0 * sg1[n] - ext1 * sg0[n])) * q2) / ext;
          }
        }
        const double f = transmit * src * absc;
        radout = radout + f;
        if (pass == 1) {
          if (have_prev) finish_prev((rtot - prefix) / transmit);
          have_prev = true;
          pv_dels = dels;
          pv_t = transmit;
          pv_abs = absc;
          pv_f = f;
          for (int k = 0; k < 8; k++) {
            pv_i0[k] = fc[k];
            pv_i1[k] = fc1[k];
            pv_pts[k] = pts[k];
          }
          prefix = prefix + f;
        }
        npassed++;
        if (npassed >= A.max_intervals) {
          atomicOr(a.errflag, 16);
          return false;
        }
      } else {
        for (int n = 0; n < 8; n++) sss[n] = 0.0;
      }
      transmit = transmit * trc;
      ext1 = ext0;
      srcext1 = srcext0;
      for (int n = 0; n < 8; n++) {
        fc1[n] = fc[n];
        sg1[n] = sg0[n];
      }
    }
    if (pass == 1) {
      for (int n = 0; n < 8; n++) {
        atomic_add_f64(&A.bsrc[(size_t)pts[n] * LD + d], wgt * beta[n]);
        if (ess) atomic_add_f64(&A.beamw[pts[n]], wgt * sss[n]);
        if (jac) {
          const int r = A.rel_s[pts[n]];
          if (r >= 0) jsrc[r] += jwt * beta[n];
          const int rb = A.rel_b[pts[n]];
          if (ess && rb >= 0) jbeam[rb] += jwt * sss[n];
        }
      }
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
    if (transmit < a.transcut) {
      validrad = true;
      rend = 0.0;
    } else if (inext < 0 && zface) {
      validrad = true;
      const int8_t tab[6][4] = {{0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5},
                                {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};
      float X[4], Y[4], R[4], DR[4];
      int bp[4];
      for (int j = 0; j < 4; j++) {
        const int ip = a.g.gridptr[(size_t)ic * 8 + tab[kface][j]];
        bp[j] = ip;
        X[j] = P[ip * 3];
        Y[j] = P[ip * 3 + 1];
        DR[j] = 0.0f;
        if ((float)mu2 < 0.0f) {
          R[j] = topv;
        } else {
          const int b = a.bot_index[ip];
          if (b < 0) atomicOr(a.errflag, 2);
          R[j] = b < 0 ? 0.0f : a.bcrad_bot[b];
          DR[j] = A.opi_gndalbedo * A.dirflux[ip];
        }
      }
      const double ub = (X[1] - X[0] > 0.0f) ? (xn - (double)X[0]) / (double)(X[1] - X[0]) : 0.0;
      const double vb = (Y[2] - Y[0] > 0.0f) ? (yn - (double)Y[0]) / (double)(Y[2] - Y[0]) : 0.0;
      const double bi[4] = {(1 - ub) * (1 - vb), ub * (1 - vb), (1 - ub) * vb, ub * vb};
      const float radbnd = (float)(bi[0] * R[0] + bi[1] * R[1] + bi[2] * R[2] + bi[3] * R[3]);
      radout = radout + transmit * (double)radbnd;
      rend = (double)radbnd;
      if (pass == 1 && ess)
        for (int j = 0; j < 4; j++) {
          atomic_add_f64(&A.beamw[bp[j]], wgt * transmit * bi[j] * (double)DR[j]);
          if (jac) {
            const int rb = A.rel_b[bp[j]];
            if (rb >= 0) jbeam[rb] += jwt * transmit * bi[j] * (double)DR[j];
          }
        }
    } else {
      icell = inext;
    }
    xe = xn;
    ye = yn;
    ze = zn;
  }
  if (pass == 1 && have_prev) finish_prev(rend);
  return true;
}

__global__ __launch_bounds__(128) void k_adjoint_rays(AdjArgs A) {
  const int ir = blockIdx.x * blockDim.x + threadIdx.x;
  if (ir >= A.r.nrays) return;
  if (A.adj_weight[ir] == 0.0 && (A.jac_weight == nullptr || A.jac_weight[ir] == 0.0)) return;
  double rtot = 0.0, r1 = 0.0;
  if (!adjoint_ray(A, ir, 0, 0.0, rtot)) return;
  adjoint_ray(A, ir, 1, rtot, r1);
}

// Contraction of the per-(point, direction) weights with GRAD8
// (COMPUTE_SOURCE_GRAD_1CELL gradient part, shdomsub4.f:1786-2013), for a
// solar source with delta-M and one Stokes component.
template <typename Emit>
__device__ void grad8_point(const GradPointArgs& a, int p, int d, Emit&& emit) {
  const int ML = a.ml, nl1 = ML + 1;
  float ry[AT3D_MAX_L + 1];
  for (int l = 0; l <= ML; l++) ry[l] = 0.0f;
  const float* Rr = a.R + (size_t)p * a.ld;
  const float* Yr = a.Yd + (size_t)d * a.ld;
  const int nr = a.nr[p];
  for (int j = 0; j < nr; j++) ry[a.lofj[j]] += Rr[j] * Yr[j];
  const float dirf = a.dirflux[p];
  const float secmu0 = a.secmu0;
  const float* ssd = a.ssd + (size_t)d * a.numphase;
  const float* dssd = a.dssd + (size_t)d * a.dnumphase;
  for (int idr = 0; idr < a.numder; idr++) {
    const int ipa = a.partder[idr];
    const size_t sp = (size_t)ipa * a.npts + p;
    // SINGSCATJ on the RTE grid
    float singscatj = 0.0f;
    for (int q = a.sj_off[sp]; q < a.sj_off[sp + 1]; q++) singscatj += a.sj_coef[q] * ssd[a.sj_iph[q]];
    float sourcet;
    if (a.npart == 1) {
      const float alb = a.albedo[sp];
      sourcet = alb > 1e-8f ? a.srcpre[(size_t)p * a.ldd + d] / alb : 0.0f;
    } else {
      sourcet = 0.0f;
      if (a.scatj_ok[sp]) {
        const float* ls = a.legt_s + sp * nl1;
        for (int l = 0; l <= ML; l++) sourcet += ls[l] * ry[l];
        sourcet = sourcet + dirf * singscatj * secmu0 / (1 - a.f_s[sp]);
      }
    }
    sourcet = fmaxf(0.0f, sourcet);
    const float* lt = a.legt_u + sp * nl1;
    const float albp_rte = a.albedo[sp];
    for (int nb = 0; nb < 8; nb++) {
      const float xi = a.optinterpwt[(size_t)p * 8 + nb];
      if (xi < 1e-7f) continue;
      const int ib = a.interpptr[(size_t)p * 8 + nb];
      const size_t pb = (size_t)ipa * a.maxpg + ib, db = (size_t)idr * a.maxpg + ib;
      const float ep = a.extinctp[pb], ap = a.albedop[pb];
      const float dext = a.dext[db], dalb = a.dalb[db];
      const float* lp = a.legp + pb * nl1;
      const float* dlp = a.dlegp + db * nl1;
      const size_t o8 = ((size_t)idr * a.npts + p) * 8 + nb;
      const float dfj = a.dfj[o8], dalbm = a.dalbm[o8];
      float dsource = 0.0f;
      for (int l = 0; l <= ML; l++) {
        const float diff = lp[l] - lt[l];
        const float dlegt = dext * diff * ap + dalb * diff * ep + dlp[l] * ep * ap + (lt[l] - 1) * dfj;
        dsource += dlegt * ry[l];
      }
      float g8 = xi * (sourcet * (albp_rte * a.dextm[db] + dalbm) + dsource);
      const int iphp = a.iphasep[pb] - 1;
      const float singscatp = a.phasewtp[pb] * ssd[iphp];
      const float dsingscatp = a.doexact[idr] ? a.phasewtp[pb] * dssd[a.diphasep[db]]
                                              : a.dphasewtp[db] * ssd[iphp];
      g8 = g8 + dirf * secmu0 * xi *
                    (singscatj * dfj + dsingscatp * ep * ap + dalb * (singscatp - singscatj) * ep +
                     dext * (singscatp - singscatj) * ap);
      emit(idr, nb, db, g8);
    }
  }
}

// GRAD8 of (point, direction) times the adjoint source weight, into the gradient.
__global__ void k_grad_points(GradPointArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.npts * a.ndir) return;
  const int p = i / a.ndir, d = i - p * a.ndir;
  const double b = a.bsrc[(size_t)p * a.ldd + d];
  if (b == 0.0) return;
  grad8_point(a, p, d, [&](int, int, size_t db, float g8) { atomic_add_f64(&a.grad[db], b * (double)g8); });
}

// GRAD8 terms of every Jacobian source slot in every ray direction.
__global__ void k_jac_g8(GradPointArgs a, JacArgs j, float* g8t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= j.nrel_s * a.ndir) return;
  const int r = i / a.ndir, d = i - r * a.ndir;
  float* o = g8t + (size_t)i * a.numder * 8;
  for (int q = 0; q < a.numder * 8; q++) o[q] = 0.0f;
  grad8_point(a, j.p_of_rs[r], d, [&](int idr, int nb, size_t, float g8) { o[idr * 8 + nb] = g8; });
}

// JACOBIAN(:, idr, ji, pixel) = RAYGRAD_PIXEL at the column's property point
// (LEVISAPPROX_GRADIENT shdomsub4.f:378-467): per ray of the pixel, the source
// weights x GRAD8, the radiance weights x DEXTM x OPTINTERPWT and minus DEXTM x
// the sun-path weights x the beam weights.  One thread per (pixel, column, derivative).
__global__ void k_jacobian(JacArgs j) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)j.npix * j.njac * j.numder) return;
  const int idr = (int)(i % j.numder);
  const int ji = (int)((i / j.numder) % j.njac);
  const int pix = (int)(i / ((int64_t)j.numder * j.njac));
  const int ib = j.jac_ib[ji];
  const float dm = j.dextm[(size_t)idr * j.maxpg + ib];
  double acc = 0.0;
  for (int ray = j.ray_off[pix]; ray < j.ray_off[pix + 1]; ray++) {
    const int d = j.dir_of_ray[ray];
    const double* js = j.jsrc + (size_t)ray * j.nrel_s;
    const double* jr = j.jrad + (size_t)ray * j.nrel_s;
    for (int q = j.js_off[ji]; q < j.js_off[ji + 1]; q++) {
      const int r = j.js_r[q], nb = j.js_nb[q];
      const float g8 = j.g8t[(((size_t)r * j.ndir + d) * j.numder + idr) * 8 + nb];
      const float xi = j.optinterpwt[(size_t)j.p_of_rs[r] * 8 + nb];
      acc = acc + js[r] * (double)g8 + (double)(dm * xi) * jr[r];
    }
    if (j.ess) {
      const double* jb = j.jbeam + (size_t)ray * j.nrel_b;
      double s = 0.0;
      for (int q = j.jb_off[ji]; q < j.jb_off[ji + 1]; q++) s = s + (double)j.jb_w[q] * jb[j.jb_r[q]];
      acc = acc - (double)dm * s;
    }
  }
  j.jacobian[i] = (float)acc;
}

// Radiance term: GRADOUT(ib) += DEXTM(ib) * OPTINTERPWT * sum of the adjoint
// radiance weights at the point (COMPUTE_RADIANCE_DERIVATIVE_ADJOINT).
__global__ void k_grad_radiance(GradPointArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.npts) return;
  const double e = a.erad[p];
  if (e == 0.0) return;
  for (int nb = 0; nb < 8; nb++) {
    const int ib = a.interpptr[(size_t)p * 8 + nb];
    const float xi = a.optinterpwt[(size_t)p * 8 + nb];
    for (int idr = 0; idr < a.numder; idr++) {
      const size_t db = (size_t)idr * a.maxpg + ib;
      const float eg = a.dextm[db] * xi;
      atomic_add_f64(&a.grad[db], (double)eg * e);
    }
  }
}

// Direct-beam term (COMPUTE_DIRECT_BEAM_DERIV_ADJOINT): per property point,
// minus DEXTM times the sun-path weights of every grid point whose beam
// crosses it, times that point's beam weight.  One thread per property point.
__global__ void k_grad_beam(GradPointArgs a) {
  const int ib = blockIdx.x * blockDim.x + threadIdx.x;
  if (ib >= a.maxpg) return;
  double s = 0.0;
  for (int e = a.bp_off[ib]; e < a.bp_off[ib + 1]; e++) s += (double)a.bp_w[e] * a.beamw[a.bp_ip[e]];
  if (s == 0.0) return;
  for (int idr = 0; idr < a.numder; idr++) {
    const size_t db = (size_t)idr * a.maxpg + ib;
    a.grad[db] = a.grad[db] - (double)a.dextm[db] * s;
  }
}

void launch_dir_source_grad(hipStream_t st, const float* raw, int ldd, int ndir, const float* ext,
                            const int* ns, const int* lofj, const float* tms, const float* G,
                            const int32_t* ss_off, const int32_t* ss_iph, const float* ss_coef,
                            const float* ssd, int numphase, int ml, int npts, int single_scatter,
                            float* dsrc, float* srcpre, float* ss8) {
  const int n = npts * ndir;
  k_dir_source_grad<<<(n + 255) / 256, 256, 0, st>>>(raw, ldd, ndir, ext, ns, lofj, tms, G, ss_off, ss_iph,
                                                      ss_coef, ssd, numphase, ml, npts, single_scatter, dsrc,
                                                      srcpre, ss8);
}

void launch_adjoint_rays(hipStream_t st, const AdjArgs& a) {
  k_adjoint_rays<<<(a.r.nrays + 127) / 128, 128, 0, st>>>(a);
}

void launch_grad_points(hipStream_t st, const GradPointArgs& a) {
  const int n = a.npts * a.ndir;
  k_grad_points<<<(n + 127) / 128, 128, 0, st>>>(a);
  k_grad_radiance<<<(a.npts + 255) / 256, 256, 0, st>>>(a);
}

void launch_grad_beam(hipStream_t st, const GradPointArgs& a) {
  k_grad_beam<<<(a.maxpg + 255) / 256, 256, 0, st>>>(a);
}

void launch_jacobian(hipStream_t st, const GradPointArgs& g, const JacArgs& j, float* g8t) {
  const int n = j.nrel_s * g.ndir;
  if (n > 0) k_jac_g8<<<(n + 127) / 128, 128, 0, st>>>(g, j, g8t);
  const int64_t m = (int64_t)j.npix * j.njac * j.numder;
  if (m > 0) k_jacobian<<<(unsigned)((m + 255) / 256), 256, 0, st>>>(j);
}

}  // namespace at3d
