// Per-point medium set-up shared by the host plan and the device kernels
// (medium.hip): the property transfer of one grid point, its delta-M
// scaling, the source and single-scatter constants and the two-stream start
// of one column.  Compiled for the host and for gfx950 from this one source
// with floating-point contraction off, so both give the same bits (the
// two-stream start calls exp, whose last bit may differ between the host's
// and the device's math library).
//
// Species-major arrays: value of species ipa at point ip is at [ipa * stride
// + ip], its NPW = 8 x NMICRO phase pointers / weights (eight corners times the
// micro-physical tables of a property point) at [(ipa * stride + ip) * NPW + q];
// phase pointers are 1-based into the prepared Legendre table.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace at3d {

// Phase tables a property point mixes at most (MAXNMICRO).
constexpr int MAX_NMICRO = 8;

// The property grid (TRILIN_INTERP_PROP's inputs).
struct PropGrid {
  const float* extinctp;   // [npart][npg]
  const float* albedop;    // [npart][npg]
  const int32_t* iphasep;  // [npart][npg][nmicro] 1-based
  const float* phasewtp;   // [npart][npg][nmicro]
  const float* zlevels;    // [npz]
  int npx, npy, npz;
  float delx, dely;
  double extmin, scatmin;
  int nmicro = 1;
  bool interp_o = false;   // INTERPMETHOD 'OO'
};

// The prepared Legendre table: coefficients / (2l+1), delta-M f subtracted
// from l <= ML (PREPARE_PROP), nl = NLEG + 1 per phase function; npw phase
// pairs per point and species.  `flat` is the whole prepared table as the
// reference lays it out, LEGEN(NSTLEG, 0:NLEG, NUMPHASE) (null: `legen`,
// NSTLEG = 1), read by RADIANCE_TRUNCATION's flat indexing.
struct LegTable {
  const float* legen;   // [numphase][nl]
  int nl, ml;
  size_t table_size;    // numphase * nl
  int npw = 8;
  const float* flat = nullptr;
  int nstleg = 1;
  size_t flat_size = 0;
  // INTERPMETHOD 'OO': the reference scales the tables themselves by
  // 1 / (1 - f) (PREPARE_PROP, shdomsub2.f:508-526), which changes the
  // rounding of RADIANCE_TRUNCATION's weights
  bool interp_o = false;
};

// SSORT (shdomsub2.f:4961-5217) with KFLAG = -2: x decreasing, y carried,
// by Singleton's quicksort (R. C. Singleton, CACM Algorithm 347, 1969), the
// reference's method, so that equal weights keep the reference's order.
template <typename T>
__host__ __device__ inline void ssort_decreasing(float* x, T* y, int n) {
  if (n < 1) return;
  float* X = x - 1;   // 1-based views; the decreasing sort is an increasing sort of -x
  T* Y = y - 1;
  for (int i = 1; i <= n; i++) X[i] = -X[i];
  int il[32], iu[32];
  int m = 1, i = 1, j = n, k = 0, l = 0, ij;
  float r = 0.375f, t;
  T ty;
  int step = 0;          // 0 select, 1 split, 2 pop, 3 short segment
  bool adjust = true;    // enter the selection through the r update
  for (;;) {
    if (step == 0) {
      if (adjust) {
        if (i == j) {
          step = 2;
          continue;
        }
        r = r <= 0.5898437f ? r + 3.90625e-2f : r - 0.21875f;
      }
      adjust = true;
      k = i;
      ij = i + (int)((float)(j - i) * r);
      t = X[ij];
      ty = Y[ij];
      if (X[i] > t) {
        X[ij] = X[i]; X[i] = t; t = X[ij];
        Y[ij] = Y[i]; Y[i] = ty; ty = Y[ij];
      }
      l = j;
      if (X[j] < t) {
        X[ij] = X[j]; X[j] = t; t = X[ij];
        Y[ij] = Y[j]; Y[j] = ty; ty = Y[ij];
        if (X[i] > t) {
          X[ij] = X[i]; X[i] = t; t = X[ij];
          Y[ij] = Y[i]; Y[i] = ty; ty = Y[ij];
        }
      }
      for (;;) {
        do l--; while (X[l] > t);
        do k++; while (X[k] < t);
        if (k > l) break;
        const float tx = X[l]; X[l] = X[k]; X[k] = tx;
        const T tt = Y[l]; Y[l] = Y[k]; Y[k] = tt;
      }
      if (l - i > j - k) {
        il[m] = i; iu[m] = l; i = k;
      } else {
        il[m] = k; iu[m] = j; j = l;
      }
      m++;
      step = 1;
    } else if (step == 2) {
      m--;
      if (m == 0) break;
      i = il[m];
      j = iu[m];
      step = 1;
    } else if (step == 1) {
      if (j - i >= 1) {
        step = 0;
        adjust = false;   // partition again without the r update
        continue;
      }
      if (i == 1) {
        step = 0;
        continue;
      }
      i--;
      step = 3;
    } else {   // straight insertion over the short segment
      for (;;) {
        i++;
        if (i == j) break;
        t = X[i + 1];
        ty = Y[i + 1];
        if (X[i] <= t) continue;
        k = i;
        do {
          X[k + 1] = X[k];
          Y[k + 1] = Y[k];
          k--;
        } while (t < X[k]);
        X[k + 1] = t;
        Y[k + 1] = ty;
      }
      step = 2;
    }
  }
  for (int q = 1; q <= n; q++) X[q] = -X[q];
}

// TRILIN_INTERP_PROP (shdom90.f90:17-346), 'ON' method, one species at one
// point: trilinear extinction, the scattering-weighted albedo and the eight
// corners' phase pointers with weights (NMICRO per corner: the property
// point's tables times its phase weights), merged by table index and sorted
// by weight as SSORT orders them, before delta-M scaling.
__host__ __device__ inline void prop_point(const PropGrid& g, int ipa, float X, float Y, float Z, float& extf,
                                           float& albf, int32_t* iph_out, float* wt_out) {
#pragma clang fp contract(off)
  const int npx = g.npx, npy = g.npy, npz = g.npz;
  const size_t npg = (size_t)npx * npy * npz;
  const float* E = g.extinctp + (size_t)ipa * npg;
  const float* A = g.albedop + (size_t)ipa * npg;
  const float* PW = g.phasewtp + (size_t)ipa * npg * g.nmicro;
  const int32_t* IP = g.iphasep + (size_t)ipa * npg * g.nmicro;
  int il = 0, iu = npz;
  while (iu - il > 1) {
    const int im = (iu + il) / 2;
    if (Z >= g.zlevels[im - 1]) il = im; else iu = im;
  }
  const int iz = il > 1 ? il : 1;
  double w = (double)(Z - g.zlevels[iz - 1]) / (double)(g.zlevels[iz] - g.zlevels[iz - 1]);
  w = w < 1.0 ? w : 1.0;
  w = w > 0.0 ? w : 0.0;
  int ix = (int)((X - 0.0f) / g.delx) + 1;
  if (fabsf(X - 0.0f - (float)npx * g.delx) < 0.01f * g.delx) ix = npx;
  const int ixp = ix % npx + 1;
  double u = (double)(X - 0.0f - g.delx * (float)(ix - 1)) / (double)g.delx;
  u = u < 1.0 ? u : 1.0;
  u = u > 0.0 ? u : 0.0;
  if (u < 1.0e-5) u = 0.0;
  if (u > 1.0 - 1.0e-5) u = 1.0;
  int iy = (int)((Y - 0.0f) / g.dely) + 1;
  if (fabsf(Y - 0.0f - (float)npy * g.dely) < 0.01f * g.dely) iy = npy;
  const int iyp = iy % npy + 1;
  double v = (double)(Y - 0.0f - g.dely * (float)(iy - 1)) / (double)g.dely;
  v = v < 1.0 ? v : 1.0;
  v = v > 0.0 ? v : 0.0;
  if (v < 1.0e-5) v = 0.0;
  if (v > 1.0 - 1.0e-5) v = 1.0;
  const double F[8] = {(1 - u) * (1 - v) * (1 - w), u * (1 - v) * (1 - w), (1 - u) * v * (1 - w),
                       u * v * (1 - w),             (1 - u) * (1 - v) * w, u * (1 - v) * w,
                       (1 - u) * v * w,             u * v * w};
  int I[8];
  I[0] = iz + npz * (iy - 1) + npz * npy * (ix - 1) - 1;
  I[1] = iz + npz * (iy - 1) + npz * npy * (ixp - 1) - 1;
  I[2] = iz + npz * (iyp - 1) + npz * npy * (ix - 1) - 1;
  I[3] = iz + npz * (iyp - 1) + npz * npy * (ixp - 1) - 1;
  for (int q = 0; q < 4; q++) I[q + 4] = I[q] + 1;
  double ext = 0.0, scat[8], scatter = 0.0;
  for (int q = 0; q < 8; q++) ext += F[q] * (double)E[I[q]];
  for (int q = 0; q < 8; q++) {
    scat[q] = F[q] * (double)E[I[q]] * (double)A[I[q]];
    scatter += scat[q];
  }
  extf = (float)ext;
  albf = ((double)extf > g.extmin) ? (float)(scatter / (double)extf) : (float)(scatter / g.extmin);
  const int nm = g.nmicro, npw = 8 * nm;
  if (g.interp_o) {
    // 'OO' (shdom90.f90:198-235, one table per property point): the corners'
    // tables in corner order, the first replaced by the table of the corner
    // with the most scattering or the one the point sits on, weights 1, 0..
    double maxscat = -1.0;
    int best = IP[I[0]];
    for (int c = 0; c < 8; c++) {
      iph_out[c] = IP[I[c]];
      wt_out[c] = 0.f;
      if (scat[c] > maxscat || fabs(F[c] - 1.0) < 0.001) {
        maxscat = scat[c];
        best = IP[I[c]];
      }
    }
    iph_out[0] = best;
    wt_out[0] = 1.f;
    return;
  }
  for (int c = 0; c < 8; c++)
    for (int m = 0; m < nm; m++) {
      const size_t src = (size_t)I[c] * nm + m;
      iph_out[c * nm + m] = IP[src];
      wt_out[c * nm + m] = (scatter >= g.scatmin) ? (float)((double)PW[src] * scat[c] / scatter)
                                                  : (float)((double)PW[src] * scat[c] / g.scatmin);
    }
  for (int q = 0; q < npw; q++)
    for (int q2 = q + 1; q2 < npw; q2++)
      if (iph_out[q] == iph_out[q2]) {
        wt_out[q] += wt_out[q2];
        wt_out[q2] = 0.f;
      }
  ssort_decreasing(wt_out, iph_out, npw);
}

// Coefficient l of the species' mixed phase function (the reference's LEGENT
// for the 'N' method): the leading table alone when its weight is at least
// 0.999, else the weighted sum over weights above 1e-5.
__host__ __device__ inline float mixed_leg(const LegTable& t, const int32_t* iph, const float* wt, int l) {
#pragma clang fp contract(off)
  if (wt[0] >= 0.999f) return t.legen[(size_t)(iph[0] - 1) * t.nl + l];
  float out = 0.f;
  for (int q = 0; q < t.npw; q++) {
    const float w = wt[q];
    if (w <= 1e-5f) continue;
    out = out + t.legen[(size_t)(iph[q] - 1) * t.nl + l] * w;
  }
  return out;
}

// The delta-M fraction f of PREPARE_PROP (shdomsub2.f:526-543): the weighted
// coefficient ML+1 over all eight weights.
__host__ __device__ inline float prepare_f(const LegTable& t, const int32_t* iph, const float* wt) {
#pragma clang fp contract(off)
  if (wt[0] >= 0.999f) return t.legen[(size_t)(iph[0] - 1) * t.nl + t.ml + 1];
  float f = 0.f;
  for (int q = 0; q < t.npw; q++) f = f + t.legen[(size_t)(iph[q] - 1) * t.nl + t.ml + 1] * wt[q];
  return f;
}

// Delta-M scaling of one species' extinction and albedo (PREPARE_PROP :547-573).
__host__ __device__ inline void deltam_scale(float f, float& ext, float& alb) {
#pragma clang fp contract(off)
  const float alb0 = alb;
  ext = (1.0f - alb0 * f) * ext;
  alb = (1.0f - f) * alb0 / (1.0f - alb0 * f);
}

// Source constants of one point (COMPUTE_SOURCE with NEWMETHOD,
// shdomsub1.f:1075-1140, and RADIANCE_TRUNCATION's weights, :1665-1735):
// the albedo-weighted, delta-M renormalised phase terms chi[0..ML], the
// mixed scattering albedo, and tc[1..ML] = sum_species W * albedo * LEGENT_l
// (tc[0] left alone).  ext / alb / iph / wt point at species 0 of this point
// with species stride `stride` (values) and 8 * stride (phase pairs).
__host__ __device__ inline void point_source_constants(const LegTable& t, int npart, bool deltam, float total_ext,
                                                       const float* ext, const float* albs, const int32_t* iph,
                                                       const float* wt, size_t stride, float* chi, float& albedo,
                                                       float* tc) {
#pragma clang fp contract(off)
  const int ML = t.ml;
  double alb = 0.0;
  for (int ipa = 0; ipa < npart; ipa++) alb = alb + (double)(ext[ipa * stride] * albs[ipa * stride]);
  for (int l = 0; l <= ML; l++) {
    float legent = 0.f;
    for (int ipa = 0; ipa < npart; ipa++) {
      const double scat = (double)(ext[ipa * stride] * albs[ipa * stride]);
      const int32_t* ip = iph + ipa * stride * t.npw;
      const float* w = wt + ipa * stride * t.npw;
      float l1 = mixed_leg(t, ip, w, l);
      if (deltam) l1 = l1 / (1 - mixed_leg(t, ip, w, ML + 1));
      legent = (float)((double)legent + scat * (double)l1);
    }
    legent = alb > 1e-10 ? (float)((double)legent / alb) : legent / (float)npart;
    chi[l] = l == 0 ? 1.0f : legent;
  }
  albedo = (float)((total_ext > 1e-10f) ? alb / (double)total_ext : 0.0);
  for (int l = 1; l <= ML; l++) {
    float rad = 0.f;
    for (int ipa = 0; ipa < npart; ipa++) {
      const float W = (total_ext == 0.0f) ? 1.0f : ext[ipa * stride] / total_ext;
      if (W == 0.0f) continue;
      const int32_t* ip = iph + ipa * stride * t.npw;
      const float* w = wt + ipa * stride * t.npw;
      float legt = mixed_leg(t, ip, w, l);
      if (deltam) {
        float F;
        if (w[0] >= 0.999f) {
          F = t.legen[(size_t)(ip[0] - 1) * t.nl + ML + 1];
        } else {
          // The reference indexes LEGEN(Q,ML+1,IPHASE(Q)) here (first index Q,
          // shdomsub1.f:1688-1689): element Q of the flat LEGEN(NSTLEG,
          // 0:NLEG, NUMPHASE) array, which with NSTLEG = 1 is coefficient
          // ML+Q (zero past the table's end).
          const float* fl = t.flat ? t.flat : t.legen;
          const size_t fsz = t.flat ? t.flat_size : t.table_size;
          F = 0.f;
          for (int q = 0; q < t.npw; q++) {
            if (w[q] <= 1e-5f) continue;
            const size_t flat = ((size_t)(ip[q] - 1) * t.nl + ML + 1) * t.nstleg + q;
            const float c = flat < fsz ? fl[flat] : 0.f;
            F = F + c * w[q];
          }
        }
        // 'OO' divides by 1 - f in the table (PREPARE_PROP), 'ON' multiplies
        // by its reciprocal here (shdomsub1.f:1690-1735)
        if (t.interp_o)
          legt = legt / (1 - F);
        else
          legt = legt * (1.0f / (1 - F));
      }
      rad = rad + W * albs[ipa * stride] * legt;
    }
    tc[l] = rad;
  }
}

// Terms 0..ML of one further phase-matrix element of one point (table te:
// a2, a3 or b1, prepared like t1): the albedo-weighted mix over species of
// the element's mixed coefficients, each divided by 1 - f of element 1 with
// delta-M (COMPUTE_SOURCE's LEGENT(2:6,L) with NEWMETHOD, NSTLEG = 6,
// shdomsub1.f:1075-1120); as chi of point_source_constants, without the
// l = 0 normalisation, which holds for element 1 only.
__host__ __device__ inline void point_element_constants(const LegTable& t1, const LegTable& te, int npart,
                                                        bool deltam, const float* ext, const float* albs,
                                                        const int32_t* iph, const float* wt, size_t stride,
                                                        float* chi_e) {
#pragma clang fp contract(off)
  const int ML = t1.ml;
  double alb = 0.0;
  for (int ipa = 0; ipa < npart; ipa++) alb = alb + (double)(ext[ipa * stride] * albs[ipa * stride]);
  for (int l = 0; l <= ML; l++) {
    float legent = 0.f;
    for (int ipa = 0; ipa < npart; ipa++) {
      const double scat = (double)(ext[ipa * stride] * albs[ipa * stride]);
      const int32_t* ip = iph + ipa * stride * t1.npw;
      const float* w = wt + ipa * stride * t1.npw;
      float l1 = mixed_leg(te, ip, w, l);
      if (deltam) l1 = l1 / (1 - mixed_leg(t1, ip, w, ML + 1));
      legent = (float)((double)legent + scat * (double)l1);
    }
    chi_e[l] = alb > 1e-10 ? (float)((double)legent / alb) : legent / (float)npart;
  }
}

// The truncated single-scatter terms of a further element (b1) per unit
// direct flux: sum_species DA * LEGENT_l / (1 - f), as tms of
// point_render_constants (COMPUTE_SOURCE_1CELL, shdomsub2.f:2960-3010).
__host__ __device__ inline void point_element_tms(const LegTable& t1, const LegTable& te, int npart,
                                                  float total_ext, const float* ext, const float* albs,
                                                  const int32_t* iph, const float* wt, size_t stride,
                                                  float secmu0, float* tms_e) {
#pragma clang fp contract(off)
  const int ML = t1.ml;
  for (int l = 0; l <= ML; l++) tms_e[l] = 0.f;
  for (int ipa = 0; ipa < npart; ipa++) {
    const float W = (total_ext == 0.0f) ? 1.0f : ext[ipa * stride] / total_ext;
    if (W == 0.0f) continue;
    const int32_t* ip = iph + ipa * stride * t1.npw;
    const float* w = wt + ipa * stride * t1.npw;
    const float f = mixed_leg(t1, ip, w, ML + 1);
    const float da = albs[ipa * stride] * secmu0 * W;
    for (int l = 0; l <= ML; l++) tms_e[l] += da * (mixed_leg(te, ip, w, l) / (1 - f));
  }
}

// Single-scatter constants of one point per unit direct flux
// (COMPUTE_SOURCE_1CELL_UNPOL's TMS terms, shdomsub2.f:3096-3175), with
// delta-M: tms[l] = sum_species DA * LEGENT_l / (1 - f) for l <= ML and the
// (phase table, DA * w / (1 - f)) pairs of the exact single-scatter term,
// DA = albedo * secmu0 * W.  Writes the pairs when ss_iph is not null and
// returns their number.
__host__ __device__ inline int point_render_constants(const LegTable& t, int npart, float total_ext,
                                                      const float* ext, const float* albs, const int32_t* iph,
                                                      const float* wt, size_t stride, float secmu0, float* tms,
                                                      int32_t* ss_iph, float* ss_coef) {
#pragma clang fp contract(off)
  const int ML = t.ml;
  if (tms)
    for (int l = 0; l <= ML; l++) tms[l] = 0.f;
  int n = 0;
  for (int ipa = 0; ipa < npart; ipa++) {
    const float W = (total_ext == 0.0f) ? 1.0f : ext[ipa * stride] / total_ext;
    if (W == 0.0f) continue;
    const int32_t* ip = iph + ipa * stride * t.npw;
    const float* w = wt + ipa * stride * t.npw;
    const float f = mixed_leg(t, ip, w, ML + 1);
    const float da = albs[ipa * stride] * secmu0 * W;
    if (tms)
      for (int l = 0; l <= ML; l++) tms[l] += da * (mixed_leg(t, ip, w, l) / (1 - f));
    if (w[0] >= 0.999f) {
      if (ss_iph) {
        ss_iph[n] = ip[0] - 1;
        ss_coef[n] = da / (1 - f);
      }
      n++;
    } else {
      for (int q = 0; q < t.npw; q++) {
        if (w[q] <= 1e-5f) continue;
        if (ss_iph) {
          ss_iph[n] = ip[q] - 1;
          ss_coef[n] = da * w[q] / (1 - f);
        }
        n++;
      }
    }
  }
  return n;
}

// Layer optical depth, albedo and asymmetry of one column of base-grid
// points for the two-stream start (INIT_RADIANCE, shdomsub2.f:614-700):
// points ip0 = column * nz + iz, layers numbered from the top.  The phase
// terms branch on species 1's leading weight for every species, as the
// reference does.
__host__ __device__ inline void eddington_layers(const LegTable& t, int npart, bool deltam, int nz,
                                                 const float* zgrid, const float* total_ext, const float* ext,
                                                 const float* albs, const int32_t* iph, const float* wt,
                                                 size_t stride, int column, float* od, float* alb_out,
                                                 float* asym) {
#pragma clang fp contract(off)
  const int ML = t.ml, nl = t.nl;
  const int nlayer = nz - 1;
  for (int iz = 1; iz <= nlayer; iz++) {
    const int L = nz - iz;
    const size_t ip0 = (size_t)(iz - 1) + (size_t)nz * column, ip1 = (size_t)iz + (size_t)nz * column;
    const float ext0 = total_ext[ip0], ext1 = total_ext[ip1];
    float scat0 = 0.f, scat1 = 0.f;
    for (int ipa = 0; ipa < npart; ipa++) {
      scat0 += albs[ipa * stride + ip0] * ext[ipa * stride + ip0];
      scat1 += albs[ipa * stride + ip1] * ext[ipa * stride + ip1];
    }
    od[L - 1] = (zgrid[iz] - zgrid[iz - 1]) * (ext0 + ext1) / 2;
    alb_out[L - 1] = (ext0 + ext1 > 0.f) ? (scat0 + scat1) / (ext0 + ext1) : 0.f;
    const int npw = t.npw;
    const float w00 = wt[ip0 * npw], w10 = wt[ip1 * npw];   // species 1's leading weights
    float g0 = 0.f, g1 = 0.f;
    for (int ipa = 0; ipa < npart; ipa++) {
      const int32_t* ph0 = iph + (ipa * stride + ip0) * npw;
      const int32_t* ph1 = iph + (ipa * stride + ip1) * npw;
      const float* wa0 = wt + (ipa * stride + ip0) * npw;
      const float* wa1 = wt + (ipa * stride + ip1) * npw;
      auto coef = [&](const int32_t* ph, const float* wa, float lead, int l) {
        if (lead >= 0.999f) return t.legen[(size_t)(ph[0] - 1) * nl + l];
        float c = 0.f;
        for (int q = 0; q < npw; q++) c = c + t.legen[(size_t)(ph[q] - 1) * nl + l] * wa[q];
        return c;
      };
      float le0 = coef(ph0, wa0, w00, 1), le1 = coef(ph1, wa1, w10, 1);
      if (deltam) {
        le0 = le0 / (1 - coef(ph0, wa0, w00, ML + 1));
        le1 = le1 / (1 - coef(ph1, wa1, w10, ML + 1));
      }
      g0 += albs[ipa * stride + ip0] * ext[ipa * stride + ip0] * le0;
      g1 += albs[ipa * stride + ip1] * ext[ipa * stride + ip1] * le1;
    }
    asym[L - 1] = (scat0 + scat1 > 0.f) ? (g0 + g1) / (scat0 + scat1) : 0.f;
  }
}

// TRIDIAG (shdomsub2.f:997-1054): a tridiagonal system with row
// interchanges, 1-based arrays of n+1 entries; false when singular.
__host__ __device__ inline bool tridiag_solve(int n, double* lower, double* diag, double* upper, double* rhs) {
#pragma clang fp contract(off)
  if (n == 1) {
    if (diag[1] == 0.0) return false;
    rhs[1] = rhs[1] / diag[1];
  }
  lower[1] = diag[1];
  diag[1] = upper[1];
  upper[1] = 0.0;
  upper[n] = 0.0;
  for (int k = 1; k <= n - 1; k++) {
    if (fabs(lower[k + 1]) >= fabs(lower[k])) {
      double x;
      x = lower[k + 1]; lower[k + 1] = lower[k]; lower[k] = x;
      x = diag[k + 1]; diag[k + 1] = diag[k]; diag[k] = x;
      x = upper[k + 1]; upper[k + 1] = upper[k]; upper[k] = x;
      x = rhs[k + 1]; rhs[k + 1] = rhs[k]; rhs[k] = x;
    }
    if (lower[k] == 0.0) return false;
    const double tt = -lower[k + 1] / lower[k];
    lower[k + 1] = diag[k + 1] + tt * diag[k];
    diag[k + 1] = upper[k + 1] + tt * upper[k];
    upper[k + 1] = 0.0;
    rhs[k + 1] = rhs[k + 1] + tt * rhs[k];
  }
  if (lower[n] == 0.0) return false;
  rhs[n] = rhs[n] / lower[n];
  rhs[n - 1] = (rhs[n - 1] - diag[n - 1] * rhs[n]) / lower[n - 1];
  for (int kb = 1; kb <= n - 2; kb++) {
    const int k = n - 2 - kb + 1;
    rhs[k] = (rhs[k] - diag[k] * rhs[k + 1] - upper[k] * rhs[k + 2]) / lower[k];
  }
  return true;
}

// The delta-Eddington two-stream fluxes of one column (EDDRTF solar branch,
// shdomsub2.f:765-993): per level (top down) the downwelling and upwelling
// diffuse fluxes and the direct flux, fluxes[3 * (nlayer + 1)]; work arrays
// of 2 * nlayer + 3 doubles each.  false: negative optical depth or a
// singular system.
__host__ __device__ inline bool eddrtf_column(int nlayer, const float* od, const float* alb, const float* asym,
                                              float solarflux, float solarmu, float gndemis, float skyrad,
                                              float* fluxes, double* lower, double* upper, double* diag,
                                              double* rhs) {
#pragma clang fp contract(off)
  const int N = 2 * nlayer + 2;
  for (int k = 0; k <= N; k++) lower[k] = upper[k] = diag[k] = rhs[k] = 0.0;
  const double PI = (double)3.1415926535f;
  const double mu0 = fabs((double)solarmu);
  double tau = 0.0;
  int I = 2;
  for (int k = 0; k < 3 * (nlayer + 1); k++) fluxes[k] = 0.f;
  for (int L = 1; L <= nlayer; L++) {
    const double deltau = od[L - 1];
    double trans, reflect, sourcep, sourcem;
    if (deltau < 0.0) return false;
    if (deltau == 0.0) {
      trans = 1.0;
      reflect = 0.0;
      sourcep = 0.0;
      sourcem = 0.0;
    } else {
      const double omega = alb[L - 1], g = asym[L - 1];
      const double R = (1.0 - omega * (4.0 - 3.0 * g)) / 4.0;
      const double T = (7.0 - omega * (4.0 + 3.0 * g)) / 4.0;
      const double lambda = sqrt(3.0 * (1.0 - omega) * (1.0 - omega * g));
      double D, X1 = 0, X2 = 0, exlp = 0, exlm = 0;
      if (lambda == 0.0) {
        D = 1.0 / (1.0 + T * deltau);
        trans = D;
        reflect = -R * deltau * D;
      } else {
        X1 = -R;
        X2 = lambda + T;
        exlp = exp(lambda * deltau < 75.0 ? lambda * deltau : 75.0);
        exlm = 1.0 / exlp;
        trans = 2. * lambda / (X2 * exlp + (lambda - T) * exlm);
        reflect = X1 * (exlp - exlm) * trans / (2. * lambda);
        D = 1.0 / (X2 * X2 * exlp - X1 * X1 * exlm);
      }
      fluxes[3 * (L - 1) + 2] = (float)((double)solarflux * exp(-tau / mu0));
      const double ds = 1.0 / (lambda * lambda - 1.0 / (mu0 * mu0));
      const double b1 = 0.5 * omega * ((double)solarflux / mu0) * exp(-tau / mu0) * ds;
      const double b2 = 0.5 * omega * ((double)solarflux / mu0) * exp(-(tau + deltau) / mu0) * ds;
      const double solpp = 1.0 + 1.5 * g * mu0, solpm = -1.0 + 1.5 * g * mu0;
      const double radp1p = ((T + 1.0 / mu0) * solpp + R * solpm) * b1;
      const double radp2m = ((-T + 1.0 / mu0) * solpm - R * solpp) * b2;
      const double radp2p = ((T + 1.0 / mu0) * solpp + R * solpm) * b2;
      const double radp1m = ((-T + 1.0 / mu0) * solpm - R * solpp) * b1;
      if (lambda == 0.0) {
        const double A = (R * deltau * radp1p - radp2m) * D;
        const double B = -(R * radp1p + T * radp2m) * D;
        sourcep = (B - T * (A + B * deltau)) / R + radp2p;
        sourcem = A + radp1m;
      } else {
        const double cp = (X1 * exlm * radp1p - X2 * radp2m) * D;
        const double cm = (-X2 * exlp * radp1p + X1 * radp2m) * D;
        sourcep = X1 * cp * exlp + X2 * cm * exlm + radp2p;
        sourcem = X2 * cp + X1 * cm + radp1m;
      }
      tau = tau + deltau;
    }
    diag[I] = -reflect;
    diag[I + 1] = -reflect;
    lower[I] = 1.0;
    lower[I + 1] = -trans;
    upper[I] = -trans;
    upper[I + 1] = 1.0;
    rhs[I] = sourcem;
    rhs[I + 1] = sourcep;
    I += 2;
  }
  fluxes[3 * nlayer + 2] = (float)((double)solarflux * exp(-tau / mu0));
  double gndflux = (1.0 - (double)gndemis) * (double)solarflux * exp(-tau / mu0);
  const double skyflux = PI * (double)skyrad;
  gndflux = gndflux + 0.0;   // no surface emission source
  rhs[1] = skyflux;
  diag[1] = 0.0;
  upper[1] = 1.0;
  diag[N] = -(1.0 - (double)gndemis);
  lower[N] = 1.0;
  rhs[N] = gndflux;
  if (!tridiag_solve(N, lower, diag, upper, rhs)) return false;
  for (int L = 1, i = 1; L <= nlayer + 1; L++, i += 2) {
    fluxes[3 * (L - 1) + 0] = (float)rhs[i];
    fluxes[3 * (L - 1) + 1] = (float)rhs[i + 1];
  }
  return true;
}

// The first radiance terms of one column's points from its two-stream
// fluxes (INIT_RADIANCE, shdomsub2.f:740-758): I0 = C0 (F_down + F_up),
// I1 = C1 (F_down - F_up), radiance0[ip][4].
__host__ __device__ inline void column_radiance(int nz, int column, const float* fl, float* radiance0) {
#pragma clang fp contract(off)
  const float PI = acosf(-1.0f);
  const float C0 = sqrtf(1.0f / PI), C1 = sqrtf(3.0f / (4 * PI));
  for (int iz = 1; iz <= nz; iz++) {
    const int L = nz + 1 - iz;
    const size_t ip = (size_t)(iz - 1) + (size_t)nz * column;
    radiance0[ip * 4 + 0] = C0 * (fl[3 * (L - 1) + 0] + fl[3 * (L - 1) + 1]);
    radiance0[ip * 4 + 2] = C1 * (fl[3 * (L - 1) + 0] - fl[3 * (L - 1) + 1]);
  }
}

}  // namespace at3d
