// Host setup of one SHDOM solve.
//
// Each function restates one piece of the reference's Fortran setup path and
// cites the routine it follows.  Arithmetic keeps the reference's mixture of
// single-precision (REAL) and double-precision (DOUBLE PRECISION) steps where
// that decides rounding of stored values, so the GPU starts from the same
// initial state as the reference.
#include "host_plan.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <thread>
#include <mutex>
#include <list>

namespace at3d {

namespace {

inline bool btest(int v, int bit) { return (v >> bit) & 1; }

// ---------------------------------------------------------------------------
// Base grid: NEW_GRIDS (shdomsub2.f:16-81) with GRIDTYPE 'P' and
// INIT_CELL_STRUCTURE (shdomsub2.f:84-309) for a base grid without a tree.
// ---------------------------------------------------------------------------
void make_grid(Plan& p) {
  const at3d_scene_desc& d = p.d;
  p.xgrid.resize(p.nx + 1);
  p.ygrid.resize(p.ny + 1);
  for (int ix = 0; ix <= p.nx; ix++) {
    float num = d.delx * (float)d.npx;
    p.xgrid[ix] = 0.0f + (float)ix * num / (float)p.nx;
  }
  for (int iy = 0; iy <= p.ny; iy++) {
    float num = d.dely * (float)d.npy;
    p.ygrid[iy] = 0.0f + (float)iy * num / (float)p.ny;
  }
  p.xgrid.resize(p.nx1);
  p.ygrid.resize(p.ny1);
  p.zgrid = p.zlevels;
  p.npts = p.nx1 * p.ny1 * p.nz;
  p.gridpos.resize((size_t)p.npts * 3);
  for (int ix = 0; ix < p.nx1; ix++)
    for (int iy = 0; iy < p.ny1; iy++)
      for (int iz = 0; iz < p.nz; iz++) {
        size_t i = iz + (size_t)p.nz * (iy + (size_t)p.ny1 * ix);
        p.gridpos[i * 3 + 0] = p.xgrid[ix];
        p.gridpos[i * 3 + 1] = p.ygrid[iy];
        p.gridpos[i * 3 + 2] = p.zgrid[iz];
      }
  const int bc = p.bcflag, ip = p.ipflag;
  int nxc = p.nx, nyc = p.ny;
  if (btest(bc, 0)) nxc = p.nx + 1;
  if (btest(bc, 2) && !btest(ip, 0)) nxc = p.nx - 1;
  if (btest(bc, 1)) nyc = p.ny + 1;
  if (btest(bc, 3) && !btest(ip, 1)) nyc = p.ny - 1;
  p.nxc = nxc;
  p.nyc = nyc;
  const int nzc = p.nz - 1;
  p.ncells = nxc * nyc * nzc;
  p.gridptr.assign((size_t)p.ncells * 8, 0);
  p.neighptr.assign((size_t)p.ncells * 6, -1);
  p.cellflags.assign(p.ncells, 0);
  const int NZ = p.nz, NY1 = p.ny1;
  for (int ix = 1; ix <= nxc; ix++) {
    int ix0, ix1;
    if (btest(ip, 0)) {
      ix0 = ix1 = ix;
    } else if (btest(bc, 0)) {
      ix0 = std::max(1, ix - 1);
      ix1 = std::min(p.nx, ix);
    } else {
      ix0 = ix;
      ix1 = ix + 1;
    }
    for (int iy = 1; iy <= nyc; iy++) {
      int iy0, iy1;
      if (btest(ip, 1)) {
        iy0 = iy1 = iy;
      } else if (btest(bc, 1)) {
        iy0 = std::max(1, iy - 1);
        iy1 = std::min(p.ny, iy);
      } else {
        iy0 = iy;
        iy1 = iy + 1;
      }
      for (int iz = 1; iz <= nzc; iz++) {
        const int I = iz + nzc * (iy - 1) + nzc * nyc * (ix - 1);  // 1-based cell
        int* gp = &p.gridptr[(size_t)(I - 1) * 8];
        auto pt = [&](int z, int y, int x) { return (z + NZ * (y - 1) + NZ * NY1 * (x - 1)) - 1; };
        gp[0] = pt(iz, iy0, ix0);
        gp[1] = pt(iz, iy0, ix1);
        gp[2] = pt(iz, iy1, ix0);
        gp[3] = pt(iz, iy1, ix1);
        gp[4] = pt(iz + 1, iy0, ix0);
        gp[5] = pt(iz + 1, iy0, ix1);
        gp[6] = pt(iz + 1, iy1, ix0);
        gp[7] = pt(iz + 1, iy1, ix1);
        int8_t fl = 0;
        int n[6];  // 1-based neighbour (0 = none), as in the reference
        if (btest(ip, 0)) {
          n[0] = I;
          n[1] = I;
          fl |= 1;
        } else if (btest(bc, 0)) {
          if (ix == 1) {
            n[0] = I;
            n[1] = I + nzc * nyc;
            fl |= 1;
          } else if (ix == nxc) {
            n[0] = I - nzc * nyc;
            n[1] = I;
            fl |= 1;
          } else {
            n[0] = I - nzc * nyc;
            n[1] = I + nzc * nyc;
          }
        } else if (btest(bc, 2)) {
          n[0] = (ix == 1) ? 0 : I - nzc * nyc;
          n[1] = (ix == nxc) ? 0 : I + nzc * nyc;
        } else {
          n[0] = (ix == 1) ? I + (nxc - 1) * nzc * nyc : I - nzc * nyc;
          n[1] = (ix == p.nx) ? I - (nxc - 1) * nzc * nyc : I + nzc * nyc;
        }
        if (btest(ip, 1)) {
          n[2] = I;
          n[3] = I;
          fl |= 2;
        } else if (btest(bc, 1)) {
          if (iy == 1) {
            n[2] = I;
            n[3] = I + nzc;
            fl |= 2;
          } else if (iy == nyc) {
            n[2] = I - nzc;
            n[3] = I;
            fl |= 2;
          } else {
            n[2] = I - nzc;
            n[3] = I + nzc;
          }
        } else if (btest(bc, 3)) {
          n[2] = (iy == 1) ? 0 : I - nzc;
          n[3] = (iy == nyc) ? 0 : I + nzc;
        } else {
          n[2] = (iy == 1) ? I + (nyc - 1) * nzc : I - nzc;
          n[3] = (iy == p.ny) ? I - (nyc - 1) * nzc : I + nzc;
        }
        n[4] = (iz == 1) ? 0 : I - 1;
        n[5] = (iz == nzc) ? 0 : I + 1;
        for (int f = 0; f < 6; f++) p.neighptr[(size_t)(I - 1) * 6 + f] = n[f] - 1;
        p.cellflags[I - 1] = fl;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Gauss-Legendre quadrature (GAUSQUADS shdomsub2.f:4874-4910) and the reduced
// Gaussian ordinate set (MAKE_ANGLE_SET shdomsub2.f:1061-1144).
// ---------------------------------------------------------------------------
void gausquads(int n, std::vector<float>& xa, std::vector<float>& wt) {
  const int k = n / 2;
  xa.assign(n, 0.f);
  wt.assign(n, 0.f);
  for (int j = 1; j <= k; j++) {
    float arg = 3.141592654f * ((float)j - 0.25f) / ((float)n + 0.5f);
    double x = (double)cosf(arg), xp, pl, pl1, pl2, dpl;
    int it = 0;
    do {
      pl1 = 1;
      pl = x;
      for (int l = 2; l <= n; l++) {
        pl2 = pl1;
        pl1 = pl;
        pl = ((2 * l - 1) * x * pl1 - (l - 1) * pl2) / l;
      }
      dpl = n * (x * pl - pl1) / (x * x - 1);
      xp = x;
      x = xp - pl / dpl;
      it++;
    } while (fabs(x - xp) > 3.0e-7 && it < 10);
    xa[j - 1] = (float)(-x);
    xa[j - 1 + k] = (float)x;
    wt[j - 1] = (float)(2.0 / ((1.0 - x * x) * dpl * dpl));
    wt[j - 1 + k] = wt[j - 1];
  }
}

// Smallest FFT-friendly azimuth count >= n used by the reduced-Gauss set: the
// values of the reference's GOODNFFT table (MAKE_ANGLE_SET shdomsub2.f:1090-1114).
int good_nfft(int n) {
  if (n <= 14) return n;
  static const int good[] = {16, 18, 20, 24, 25, 27, 30, 32, 36, 40, 45, 48, 50, 54, 60, 64, 72,
                             80, 81, 90, 96, 100, 108, 120, 128, 144, 160, 180, 192, 200, 216, 240, 256};
  for (int g : good)
    if (g >= n) return g;
  return -1;
}

int make_angle_set(Plan& p, std::string& err) {
  const int nmu = p.nmu;
  if (p.d.angle_set != 1 && p.d.angle_set != 2) {
    err = "angle_set must be 1 (gaussian) or 2 (reduced gaussian)";
    return AT3D_ERR_ARG;
  }
  if (p.nphi0max > 256 && p.d.angle_set == 2) {
    err = "MAKE_ANGLE_SET: exceeded GOODNFFT";
    return AT3D_ERR_ARG;
  }
  gausquads(nmu, p.mu, p.wtmu);
  p.nphi0.assign(nmu, 0);
  p.phi.assign((size_t)nmu * p.nphi0max, 0.f);
  p.wtdo.assign((size_t)nmu * p.nphi0max, 0.f);
  p.nang = 0;
  for (int j = 0; j < nmu; j++) {
    int n0;
    if (p.d.angle_set == 1) {
      n0 = p.nphi0max;
    } else {
      n0 = (int)(0.9f + (float)p.nphi0max * sqrtf(1.0f - p.mu[j] * p.mu[j]));
      n0 = std::min(p.nphi0max, good_nfft(n0));
    }
    p.nphi0[j] = n0;
    float delphi = 2.0f * acosf(-1.0f) / (float)n0;
    for (int k = 0; k < n0; k++) {
      p.phi[(size_t)j * p.nphi0max + k] = (float)k * delphi;
      p.wtdo[(size_t)j * p.nphi0max + k] = delphi * p.wtmu[j];
    }
    p.nang += n0;
  }
  return AT3D_OK;
}

// ---------------------------------------------------------------------------
// Normalised Wigner d^l_{m0} (WIGNERFCT02P2M_NORMALIZED DM0 branch,
// shdomsub2.f:4363-4400, used by PLMALL) and the unpolarised variant
// (WIGNERFCT_DM0, shdomsub2.f:4542-4578, used by YLMALL_UNPOL).
// ---------------------------------------------------------------------------
double dmm1_n0(double x, int m, int m1) {  // DMM1_N0 shdomsub2.f:4452-4487
  if (m == m1) return pow((1.0 + x) / 2.0, m);
  double cmm1 = (m1 > m) ? 1.0 : ((m - m1) % 2 ? -1.0 : 1.0);
  int maxm = std::max(m, m1), minm = std::min(m, m1);
  double prod = 1.0;
  for (int pp = 1; pp <= maxm - minm; pp++) prod *= sqrt((double)(m + m1 + pp) / (double)pp);
  double f = sqrt((1.0 - x) / 2.0);
  double r = cmm1 * prod * pow(f, maxm - minm);
  f = sqrt((1.0 + x) / 2.0);
  return r * pow(f, maxm + minm);
}

void wigner_dm0_plm(double x, int nrank, int m, std::vector<double>& dm0) {
  dm0.assign(nrank + 1, 0.0);
  if (m > nrank) return;
  if (m == 0) {
    dm0[0] = 1.0;
    if (nrank >= 1) dm0[1] = x;
    for (int n = 1; n <= nrank - 1; n++) {
      double f1 = (double)(2 * n + 1) * x / (double)(n + 1);
      double f2 = (double)n / (double)(n + 1);
      dm0[n + 1] = f1 * dm0[n] - f2 * dm0[n - 1];
    }
  } else {
    dm0[m] = dmm1_n0(x, m, 0);
    for (int n = m; n <= nrank - 1; n++) {
      double f1 = (double)(n * (n + 1)) * x;
      f1 = f1 / sqrt((double)((n + 1) * (n + 1) - m * m));
      f1 = f1 / (double)(n + 1);
      f1 = f1 * (double)(2 * n + 1) / (double)n;
      double f2 = sqrt((double)(n * n - m * m)) * (double)n;
      f2 = f2 / sqrt((double)((n + 1) * (n + 1) - m * m));
      f2 = f2 / (double)(n + 1);
      f2 = f2 * (double)(n + 1) / (double)n;
      dm0[n + 1] = f1 * dm0[n] - f2 * dm0[n - 1];
    }
  }
  for (int n = 0; n <= nrank; n++) dm0[n] = sqrt(n + 0.5) * dm0[n];
}

void wigner_dm0_ylm(double x, int nrank, int m, std::vector<double>& dm0) {
  dm0.assign(nrank + 1, 0.0);
  if (m > nrank) return;
  if (m == 0) {
    dm0[0] = 1.0;
    if (nrank >= 1) dm0[1] = x;
    for (int n = 1; n <= nrank - 1; n++)
      dm0[n + 1] = ((2 * n + 1) * x * dm0[n] - n * dm0[n - 1]) / (n + 1);
  } else {
    double cm = (m % 2) ? -1.0 : 1.0, prod = 1.0;
    for (int pp = 1; pp <= m; pp++) prod *= sqrt((m + pp) / (double)pp);
    dm0[m] = cm * prod * pow(0.5 * sqrt((1.0 - x) * (1.0 + x)), m);
    for (int n = m; n <= nrank - 1; n++)
      dm0[n + 1] = ((2 * n + 1) * x * dm0[n] - sqrt((double)(n * n - m * m)) * dm0[n - 1]) /
                   sqrt((double)((n + 1) * (n + 1) - m * m));
  }
  for (int n = 0; n <= nrank; n++) dm0[n] = sqrt(n + 0.5) * dm0[n];
}

inline int sh_index(int l, int m, int mm) {  // 0-based J for (l,m)
  return (l <= mm) ? l * (l + 1) + m : (2 * mm + 1) * l - mm * mm + m;
}

// The normalised Wigner functions of PLMALL / YLMALL for NSTLEG = 6
// (WIGNERFCT02P2M_NORMALIZED, shdomsub2.f:4363-4450): d^l_{m0}(mu) as
// wigner_dm0_plm, and d^l_{m2}(mu), (-1)^(l+m) d^l_{m2}(-mu), each times
// sqrt(l + 1/2).
void wigner_d2pm(double x, int nrank, int m, std::vector<double>& dm2p, std::vector<double>& dm2m) {
  dm2p.assign(nrank + 1, 0.0);
  dm2m.assign(nrank + 1, 0.0);
  if (m > nrank || nrank < 2) return;
  wigner_d(x, nrank, m, 2, dm2p);
  wigner_d(-x, nrank, m, 2, dm2m);
  for (int n = 0; n <= nrank; n++) {
    if ((n + m) % 2) dm2m[n] = -dm2m[n];
    dm2p[n] = sqrt(n + 0.5) * dm2p[n];
    dm2m[n] = sqrt(n + 0.5) * dm2m[n];
  }
}

// The zenith part of the six-element real generalised spherical harmonics
// (PLMALL, shdomsub2.f:4650-4752), [nlm][6] in single precision: with
// P1 = d^l_{m0}, P2 = -(D2P + D2M) / 2, P3 = -(D2P - D2M) / 2 (times
// 1 / sqrt(2 pi)), elements (P1, P2, P2, P1, P3, P3) for m >= 0 and
// (P1, P2, -P2, -P1, s P3, -s P3) for m < 0, s = -1 for the transposed set
// the analysis uses, +1 otherwise.
void plm_pol(double x, int ml, int mm, bool transpose, std::vector<float>& prc) {
  const int nlm = (2 * mm + 1) * (ml + 1) - mm * (mm + 1);
  prc.assign((size_t)nlm * 6, 0.f);
  const double fct = 1.0 / sqrt(2.0 * acos(-1.0));
  const double sgn = transpose ? -1.0 : 1.0;
  std::vector<double> dm0, dp, dm;
  for (int m = 0; m <= mm; m++) {
    wigner_dm0_plm(x, ml, m, dm0);
    wigner_d2pm(x, ml, m, dp, dm);
    for (int l = m; l <= ml; l++) {
      const float p1 = (float)(fct * dm0[l]);
      const float p2 = (float)(-0.5 * fct * (dp[l] + dm[l]));
      const float p3 = (float)(-0.5 * fct * (dp[l] - dm[l]));
      float* a = &prc[(size_t)sh_index(l, m, mm) * 6];
      a[0] = p1; a[1] = p2; a[2] = p2; a[3] = p1; a[4] = p3; a[5] = p3;
      if (m > 0) {
        float* b = &prc[(size_t)sh_index(l, -m, mm) * 6];
        b[0] = p1; b[1] = p2; b[2] = -p2; b[3] = -p1;
        b[4] = (float)(sgn * p3); b[5] = (float)(-sgn * p3);
      }
    }
  }
}

// YLMALL_UNPOL (shdomsub2.f:4490-4539): real generalised SH of one direction.
void ylm_unpol(float mu, float phi, int ml, int mm, float* yr) {
  const double fct = 1.0 / sqrt(2.0 * acos(-1.0));
  std::vector<double> dm0;
  for (int m = 0; m <= mm; m++) {
    wigner_dm0_ylm((double)mu, ml, m, dm0);
    double cosm = 1.0, sinm = 0.0;
    if (m > 0) {
      cosm = (double)cosf((float)m * phi);
      sinm = (double)sinf((float)m * phi);
    }
    for (int l = m; l <= ml; l++) {
      double v = fct * dm0[l];
      yr[sh_index(l, m, mm)] = (float)((cosm - sinm) * v);
      yr[sh_index(l, -m, mm)] = (float)((cosm + sinm) * v);
    }
  }
}

void make_sh_tables_pol(Plan& p);
void prepare_legen_pol(Plan& p);

// Dense SH<->DO transform matrices.  SH_TO_DO_UNPOL (shdomsub1.f:2928-3036)
// sums P_lm(mu) over l per m (CMU1 = PLMALL, MAKE_SH_DO_COEF shdomsub2.f:1149-1220)
// and then synthesises the azimuth with cos/sin of m*phi (FFT or DFT, both
// identical in exact arithmetic) for |m| <= NPHI0/2-1.  DO_TO_SH_UNPOL
// (shdomsub1.f:3164-3254) is the weighted transpose with weight
// WTMU * (2 pi / NPHI0).  Both are written here as one [nang][nlm] matrix each.
void make_sh_tables(Plan& p) {
  const int ml = p.ml, mm = p.mm, nlm = p.nlm;
  p.lofj.assign(nlm, 0);
  p.mofj.assign(nlm, 0);
  for (int l = 0, j = 0; l <= ml; l++)
    for (int m = -std::min(l, mm); m <= std::min(l, mm); m++, j++) {
      p.lofj[j] = l;
      p.mofj[j] = m;
    }
  const double fct = 1.0 / sqrt(2.0 * acos(-1.0));
  p.ymat.assign((size_t)p.nang * nlm, 0.f);
  p.wmat.assign((size_t)p.nang * nlm, 0.f);
  std::vector<float> plm(nlm);
  std::vector<double> dm0;
  int k = 0;
  for (int imu = 0; imu < p.nmu; imu++) {
    for (int m = 0; m <= mm; m++) {
      wigner_dm0_plm((double)p.mu[imu], ml, m, dm0);
      for (int l = m; l <= ml; l++) {
        float v = (float)(fct * dm0[l]);
        plm[sh_index(l, m, mm)] = v;
        plm[sh_index(l, -m, mm)] = v;
      }
    }
    const int n0 = p.nphi0[imu];
    const int me = std::max(0, std::min(n0 / 2 - 1, mm));
    const double dphi = 2.0 * acos(-1.0) / n0;
    const float wmu = p.wtmu[imu];
    for (int ip = 0; ip < n0; ip++, k++) {
      const double ph = ip * dphi;
      for (int j = 0; j < nlm; j++) {
        const int m = p.mofj[j], am = std::abs(m);
        if (am > me) continue;
        double t;
        if (m == 0)
          t = 1.0;
        else if (m > 0)
          t = cos(am * ph) - sin(am * ph);
        else
          t = cos(am * ph) + sin(am * ph);
        p.ymat[(size_t)k * nlm + j] = (float)((double)plm[j] * t);
        p.wmat[(size_t)k * nlm + j] = (float)((double)plm[j] * (double)wmu * dphi * t);
      }
    }
  }
  p.ylmsun.assign(nlm, 0.f);
  ylm_unpol(p.d.solarmu, p.d.solaraz, ml, mm, p.ylmsun.data());
  if (p.nst == 3) make_sh_tables_pol(p);
}

// The polarized transforms (SH_TO_DO / DO_TO_SH with NSTOKES = 3,
// shdomsub1.f:2789-2927 and 3041-3160) as dense tables per phase-matrix
// element e: Stokes component I of an ordinate is sum_J Y1_J S^I_J, Q is
// sum_J Y2_J S^Q_J + Y5_J S^U_J, U is sum_J Y6_J S^Q_J + Y3_J S^U_J, with
// Ye_J = CMU1(e, J, mu) times the azimuth factor of the component: 1 for
// m = 0, cos(m phi) -/+ sin(m phi) for I and Q and cos(m phi) +/- sin(m phi)
// for U (upper sign m > 0), which is how SUMCS combines the +m and -m sums;
// the analysis is the weighted transpose with CMU2 (the transposed set),
// Q's azimuth factors on the elements fed by Q (2, 6) and U's on those fed
// by U (3, 5).  The same m cut-off at NPHI0/2 - 1 as the scalar tables; the
// I tables are the scalar ones.  The sun's functions (element 1 of YLMALL,
// INIT_SOLUTION's YLMSUN) come from the same zenith functions.
void make_sh_tables_pol(Plan& p) {
  const int ml = p.ml, mm = p.mm, nlm = p.nlm, nang = p.nang;
  p.ypol.assign((size_t)6 * nang * nlm, 0.f);
  p.wpol.assign((size_t)6 * nang * nlm, 0.f);
  std::vector<float> prc, prct;
  int k = 0;
  for (int imu = 0; imu < p.nmu; imu++) {
    plm_pol((double)p.mu[imu], ml, mm, false, prc);
    plm_pol((double)p.mu[imu], ml, mm, true, prct);
    const int n0 = p.nphi0[imu];
    const int me = std::max(0, std::min(n0 / 2 - 1, mm));
    const double dphi = 2.0 * acos(-1.0) / n0;
    const float wmu = p.wtmu[imu];
    for (int ip = 0; ip < n0; ip++, k++) {
      const double ph = ip * dphi;
      for (int j = 0; j < nlm; j++) {
        const int m = p.mofj[j], am = std::abs(m);
        if (am > me) continue;
        const double c = cos(am * ph), sn = sin(am * ph);
        const double fi = m == 0 ? 1.0 : (m > 0 ? c - sn : c + sn);   // I and Q
        const double fu = m == 0 ? 1.0 : (m > 0 ? c + sn : c - sn);   // U
        const double yf[6] = {fi, fi, fu, 0.0, fi, fu};
        const double wf[6] = {fi, fi, fu, 0.0, fu, fi};
        for (int e = 0; e < 6; e++) {
          if (e == 3) continue;
          const size_t o = ((size_t)e * nang + k) * nlm + j;
          p.ypol[o] = (float)((double)prc[(size_t)j * 6 + e] * yf[e]);
          p.wpol[o] = (float)((double)prct[(size_t)j * 6 + e] * (double)wmu * dphi * wf[e]);
        }
      }
    }
  }
  std::vector<float> ys((size_t)nlm * 6);
  ylm_direction_pol(p.d.solarmu, p.d.solaraz, ml, mm, ys.data());
  for (int j = 0; j < nlm; j++) p.ylmsun[j] = ys[(size_t)j * 6];
}

// ---------------------------------------------------------------------------
// Property transfer: TRILIN_INTERP_PROP (shdom90.f90:17-346, 'ON' method),
// INTERP_GRID (shdomsub2.f:313-387) and PREPARE_PROP (shdomsub2.f:479-608).
// ---------------------------------------------------------------------------
}  // namespace

// TRILIN_INTERP_PROP (shdom90.f90:17-346) for one species at one point
// (X, Y, Z): trilinear extinction, the scattering-weighted albedo and the
// phase-table entries of the eight property-grid corners with their weights
// (merged by table index and sorted by weight, 'N' interpolation), before
// delta-M scaling.
PropGrid plan_prop_grid(const Plan& p) {
  const at3d_scene_desc& d = p.d;
  return PropGrid{p.extinctp.data(), p.albedop.data(), p.iphasep.data(), p.phasewtp.data(), p.zlevels.data(),
                  d.npx, d.npy, d.npz, d.delx, d.dely, p.extmin, p.scatmin, p.nmicro, p.interp_o};
}

LegTable plan_leg_table(const Plan& p) {
  LegTable t{p.legen.data(), p.nleg + 1, p.ml, p.legen.size(), p.npw};
  t.interp_o = p.interp_o;
  if (!p.legen6flat.empty()) {
    t.flat = p.legen6flat.data();
    t.nstleg = 6;
    t.flat_size = p.legen6flat.size();
  }
  return t;
}

float sky_flux_weight(const Plan& p) {
  float skyradalb = 0.f;
  for (int imu = 0; imu < p.nmu / 2; imu++)
    for (int ip = 0; ip < p.nphi0[imu]; ip++)
      skyradalb = skyradalb + fabsf(p.mu[imu]) * p.wtdo[(size_t)imu * p.nphi0max + ip] * p.d.skyrad;
  return skyradalb;
}

void point_optics(const Plan& p, int ipa, float X, float Y, float Z, float& extf, float& albf, int32_t* iph_out,
                  float* wt_out) {
  prop_point(plan_prop_grid(p), ipa, X, Y, Z, extf, albf, iph_out, wt_out);
}

namespace {

// Tabulated phase functions divided by (2l+1) (TRILIN_INTERP_PROP's set-up)
// and, with delta-M, f = chi_{ML+1} subtracted from l <= ML (PREPARE_PROP,
// 'N' interpolation).
void prepare_legen(Plan& p) {
  const int nl = p.nleg + 1, ML = p.ml;
  p.legen.assign((size_t)p.numphase * nl, 0.f);
  for (int iph = 0; iph < p.numphase; iph++)
    for (int l = 0; l < nl; l++)
      p.legen[(size_t)iph * nl + l] = p.legenp[(size_t)iph * (p.nlegp + 1) + l] / (float)(2 * l + 1);
  if (p.d.deltam)
    for (int iph = 0; iph < p.numphase; iph++) {
      const float f = p.legen[(size_t)iph * nl + ML + 1];
      for (int l = 0; l <= ML; l++) p.legen[(size_t)iph * nl + l] = p.legen[(size_t)iph * nl + l] - f;
    }
}

// EXTMIN and SCATMIN of TRILIN_INTERP_PROP (shdom90.f90:60-62).
void medium_scalars(Plan& p) {
  const int npz = p.d.npz;
  p.extmin = (double)(1.0e-5f / ((p.zlevels[npz - 1] - p.zlevels[0]) / (float)npz));
  p.scatmin = (double)0.1f * p.extmin;
}

void interp_optics(Plan& p) {
  const int npts = p.npts, npart = p.npart;
  medium_scalars(p);
  p.extinct.assign((size_t)npart * npts, 0.f);
  p.albedo.assign((size_t)npart * npts, 0.f);
  const int npw = p.npw;
  p.iphase.assign((size_t)npart * npts * npw, 1);
  p.phaseinterpwt.assign((size_t)npart * npts * npw, 0.f);
  p.total_ext.assign(npts, 0.f);
  for (int ipa = 0; ipa < npart; ipa++) {
    // points are independent; the species loop stays outside so each point's
    // total extinction is summed in species order
    const int nchunk = 64;
    parallel_for(nchunk, [&](int c) {
    const int ip0 = (int)((int64_t)npts * c / nchunk), ip1 = (int)((int64_t)npts * (c + 1) / nchunk);
    for (int ip = ip0; ip < ip1; ip++) {
      const size_t base = ((size_t)ipa * npts + ip) * npw;
      float extf, albf;
      point_optics(p, ipa, p.gridpos[ip * 3], p.gridpos[ip * 3 + 1], p.gridpos[ip * 3 + 2], extf, albf,
                   &p.iphase[base], &p.phaseinterpwt[base]);
      p.extinct[(size_t)ipa * npts + ip] = extf;
      p.albedo[(size_t)ipa * npts + ip] = albf;
      p.total_ext[ip] = p.total_ext[ip] + extf;
    }
    });
  }
  prepare_legen(p);
  if (p.nst == 3) prepare_legen_pol(p);
  const LegTable lt = plan_leg_table(p);
  if (p.d.deltam) {
    for (int ip = 0; ip < npts; ip++) {
      float s = 0.f;
      for (int ipa = 0; ipa < npart; ipa++) s += p.extinct[(size_t)ipa * npts + ip];
      float t = p.total_ext[ip] - s;
      p.total_ext[ip] = t < 0.f ? 0.f : t;
    }
    for (int ipa = 0; ipa < npart; ipa++)
      for (int ip = 0; ip < npts; ip++) {
        const size_t base = ((size_t)ipa * npts + ip) * npw;
        const float f = prepare_f(lt, &p.iphase[base], &p.phaseinterpwt[base]);
        float& ext = p.extinct[(size_t)ipa * npts + ip];
        deltam_scale(f, ext, p.albedo[(size_t)ipa * npts + ip]);
        p.total_ext[ip] = p.total_ext[ip] + ext;
      }
  }
  p.albmax = 0.f;
  for (float a : p.albedo) p.albmax = std::max(p.albmax, a);
}

// Mixed phase function LEGENT1 of one species at one point ('N' branch of
// COMPUTE_SOURCE, shdomsub1.f:1088-1112): returns coefficients 0..nleg.
void mixed_legen(const Plan& p, int ipa, int ip, std::vector<float>& out) {
  const int nl = p.nleg + 1;
  const size_t base = ((size_t)ipa * p.npts + ip) * p.npw;
  out.assign(nl, 0.f);
  if (p.phaseinterpwt[base] >= 0.999f) {
    const float* L = &p.legen[(size_t)(p.iphase[base] - 1) * nl];
    for (int l = 0; l < nl; l++) out[l] = L[l];
  } else {
    for (int q = 0; q < p.npw; q++) {
      const float w = p.phaseinterpwt[base + q];
      if (w <= 1e-5f) continue;
      const float* L = &p.legen[(size_t)(p.iphase[base + q] - 1) * nl];
      for (int l = 0; l < nl; l++) out[l] = out[l] + L[l] * w;
    }
  }
}

// Per-point constants of the source computation: the scattering albedo and
// mixed, delta-M renormalised Legendre coefficients (COMPUTE_SOURCE NEWMETHOD,
// shdomsub1.f:1075-1140) and the radiance-truncation weights
// (RADIANCE_TRUNCATION, shdomsub1.f:1665-1735).
void source_constants(Plan& p) {
  const int ML = p.ml, npts = p.npts, npart = p.npart;
  p.chi.assign((size_t)npts * (ML + 1), 0.f);
  p.alb.assign(npts, 0.f);
  p.tc.assign((size_t)npts * (ML + 1), 0.f);
  const LegTable lt = plan_leg_table(p);
  // points are independent: chunks over the host threads
  const int nchunk = 64;
  parallel_for(nchunk, [&](int c) {
    const int i0 = (int)((int64_t)npts * c / nchunk), i1 = (int)((int64_t)npts * (c + 1) / nchunk);
    for (int i = i0; i < i1; i++)
      point_source_constants(lt, npart, p.d.deltam != 0, p.total_ext[i], &p.extinct[i], &p.albedo[i],
                             &p.iphase[(size_t)i * p.npw], &p.phaseinterpwt[(size_t)i * p.npw], (size_t)npts,
                             &p.chi[(size_t)i * (ML + 1)], p.alb[i], &p.tc[(size_t)i * (ML + 1)]);
  });
}

// The prepared tables of the phase-matrix elements (NSTLEG = 6): divided by
// (2l+1) (TRILIN_INTERP_PROP's set-up); with delta-M, element 1's f
// subtracted from l <= ML of a2, a3 and a4, b1 and b2 left alone
// (PREPARE_PROP, shdomsub2.f:526-543, 'N' interpolation).  a2, a3 and b1 as
// tables of their own, all six in the reference's interleaved order.
void prepare_legen_pol(Plan& p) {
  const int nl = p.nleg + 1, ML = p.ml, np = p.numphase;
  std::vector<float> el[6];
  for (int e = 0; e < 6; e++) {
    std::vector<float>& t = el[e];
    if (e == 0) {
      t = p.legen;
      continue;
    }
    t.assign((size_t)np * nl, 0.f);
    for (int iph = 0; iph < np; iph++)
      for (int l = 0; l < nl; l++)
        t[(size_t)iph * nl + l] = p.legenp6[((size_t)e * np + iph) * (p.nlegp + 1) + l] / (float)(2 * l + 1);
    if (p.d.deltam && e < 4)
      for (int iph = 0; iph < np; iph++) {
        const float f = p.legen[(size_t)iph * nl + ML + 1];
        for (int l = 0; l <= ML; l++) t[(size_t)iph * nl + l] = t[(size_t)iph * nl + l] - f;
      }
  }
  p.legen2 = el[1];
  p.legen3 = el[2];
  p.legen5 = el[4];
  p.legen6flat.assign((size_t)np * nl * 6, 0.f);
  for (int iph = 0; iph < np; iph++)
    for (int l = 0; l < nl; l++)
      for (int e = 0; e < 6; e++) p.legen6flat[((size_t)iph * nl + l) * 6 + e] = el[e][(size_t)iph * nl + l];
}

// Per point, the a2, a3 and b1 terms of the polarized source
// (CALC_SOURCE_PNT, shdomsub1.f:823-910, takes LEGEN elements 1, 2, 3, 5)
// and the b1 single-scatter terms of rendering, per unit direct flux.
void polarized_constants(Plan& p) {
  const int ML = p.ml, npts = p.npts, npart = p.npart, nl = p.nleg + 1;
  p.chip.assign((size_t)npts * 3 * (ML + 1), 0.f);
  p.tms5.assign((size_t)npts * (ML + 1), 0.f);
  const LegTable t1 = plan_leg_table(p);
  const LegTable te[3] = {LegTable{p.legen2.data(), nl, ML, p.legen2.size(), p.npw},
                          LegTable{p.legen3.data(), nl, ML, p.legen3.size(), p.npw},
                          LegTable{p.legen5.data(), nl, ML, p.legen5.size(), p.npw}};
  const float secmu0 = (float)(1.0 / fabs((double)p.d.solarmu));
  const int nchunk = 64;
  parallel_for(nchunk, [&](int c) {
    const int i0 = (int)((int64_t)npts * c / nchunk), i1 = (int)((int64_t)npts * (c + 1) / nchunk);
    for (int i = i0; i < i1; i++) {
      for (int k = 0; k < 3; k++)
        point_element_constants(t1, te[k], npart, p.d.deltam != 0, &p.extinct[i], &p.albedo[i],
                                &p.iphase[(size_t)i * p.npw], &p.phaseinterpwt[(size_t)i * p.npw], (size_t)npts,
                                &p.chip[((size_t)i * 3 + k) * (ML + 1)]);
      if (p.d.deltam)
        point_element_tms(t1, te[2], npart, p.total_ext[i], &p.extinct[i], &p.albedo[i],
                          &p.iphase[(size_t)i * p.npw], &p.phaseinterpwt[(size_t)i * p.npw], (size_t)npts, secmu0,
                          &p.tms5[(size_t)i * (ML + 1)]);
    }
  });
}

// ---------------------------------------------------------------------------
// Direct solar beam through the property grid: DIRECT_BEAM_PROP
// (shdom90.f90:352-867) called by MAKE_DIRECT (shdomsub2.f:393-473).
// ---------------------------------------------------------------------------
}  // namespace

void direct_beam_setup(const Plan& p, std::vector<float>& extdirp, DirectBeamGeom& b) {
  const at3d_scene_desc& d = p.d;
  const int npx = d.npx, npy = d.npy, npz = d.npz, npg = npx * npy * npz;
  const int ML = p.ml, npart = p.npart;
  extdirp.assign(npg, 0.f);
  for (int ip = 0; ip < npg; ip++) {
    float acc = 0.f;
    for (int ipa = 0; ipa < npart; ipa++) {
      double ext = p.extinctp[(size_t)ipa * npg + ip];
      double alb = p.albedop[(size_t)ipa * npg + ip];
      if (d.deltam) {
        // f over the point's phase tables, summed in double
        const int L = ML + 1;
        double F = 0.0;
        for (int q = 0; q < p.nmicro; q++) {
          const size_t pq = ((size_t)ipa * npg + ip) * p.nmicro + q;
          const int iph = p.iphasep[pq];
          float term = p.phasewtp[pq] * p.legenp[(size_t)(iph - 1) * (p.nlegp + 1) + L];
          term = term / (float)(2 * L + 1);
          F = F + (double)term;
        }
        ext = (1.0 - alb * F) * ext;
      }
      acc = (float)((double)acc + ext);
    }
    extdirp[ip] = acc;
  }
  direct_beam_geometry(p, b);
  b.ext = extdirp.data();
}

void direct_beam_geometry(const Plan& p, DirectBeamGeom& b) {
  const at3d_scene_desc& d = p.d;
  const int npx = d.npx, npy = d.npy, npz = d.npz;
  b.ipdirect = p.ipflag;
  if (btest(p.ipflag, 2)) b.ipdirect = 0;
  const double sunmu = -(double)d.solarmu;
  const double sunaz = (double)(d.solaraz + acosf(-1.0f));
  b.cx = sqrt(1.0 - sunmu * sunmu) * cos(sunaz);
  b.cy = sqrt(1.0 - sunmu * sunmu) * sin(sunaz);
  b.cz = fabs(sunmu);
  if (fabs(b.cx) > 1.0e-6) b.cxinv = 1.0 / b.cx; else { b.cx = 0.0; b.cxinv = 1.0e20; }
  if (fabs(b.cy) > 1.0e-6) b.cyinv = 1.0 / b.cy; else { b.cy = 0.0; b.cyinv = 1.0e20; }
  if (fabs(b.cz) > 1.0e-6) b.czinv = 1.0 / b.cz; else { b.cz = 0.0; b.czinv = 1.0e20; }
  b.di = b.cx >= 0 ? 1 : -1;
  b.dj = b.cy >= 0 ? 1 : -1;
  b.dk = b.cz >= 0 ? 1 : -1;
  const float zspan = p.zlevels[npz - 1] - p.zlevels[0];
  b.epsz = (double)(1.0e-6f * zspan);
  b.epss = (double)(1.0e-3f * zspan / (float)npz);
  if (!btest(b.ipdirect, 0)) b.epss = std::max(b.epss, 1.0e-4 * (double)d.delx);
  if (!btest(b.ipdirect, 1)) b.epss = std::max(b.epss, 1.0e-4 * (double)d.dely);
  b.delxd = d.delx;
  b.delyd = d.dely;
  b.epss = std::max(std::max((double)0.001f * b.delxd, (double)0.001f * b.delyd), b.epss);
  b.xdomain = b.delxd * npx;
  if (btest(p.bcflag, 2)) b.xdomain = b.delxd * (npx - 1);
  b.ydomain = b.delyd * npy;
  if (btest(p.bcflag, 3)) b.ydomain = b.delyd * (npy - 1);
  b.zl = p.zlevels.data();
  b.ext = nullptr;
  b.npx = npx;
  b.npy = npy;
  b.npz = npz;
  b.bc = p.bcflag;
  b.solarflux = (double)d.solarflux;
}

// Direct flux at one point (false when the reference would STOP).  With
// `path` the traversal also records, per property-grid cell crossed, the
// derivative of the slant optical path with respect to each of the 8 corner
// extinctions (DIRECT_BEAM_AND_PATHS_PROP, shdomsub5.f:1608-2003).
bool direct_beam_point(const DirectBeamGeom& b, float xi, float yi, float zi, float& dirflux,
                       std::vector<std::pair<int32_t, float>>* path_rec) {
  double path = 0.0;
  bool ok;
  if (path_rec)
    ok = direct_beam_trace(b, xi, yi, zi, path, [&](const int* ii, const double* seg) {
      for (int c = 0; c < 8; c++) path_rec->emplace_back(ii[c], (float)seg[c]);
    });
  else
    ok = direct_beam_trace(b, xi, yi, zi, path, NoBeamRecord{});
  if (!ok) return false;
  dirflux = (float)(b.solarflux * exp(-path));
  return true;
}

namespace {

// ---------------------------------------------------------------------------
// Eddington two-stream initial radiance: INIT_RADIANCE (shdomsub2.f:614-760),
// EDDRTF (shdomsub2.f:765-993) solar branch and TRIDIAG (shdomsub2.f:997-1054).
// ---------------------------------------------------------------------------
// The two-stream start of every column (INIT_RADIANCE with EDDRTF,
// shdomsub2.f:614-993): layers from the points' optics, the delta-Eddington
// fluxes, and the radiance's first terms I0 = C0 (F_down + F_up),
// I1 = C1 (F_down - F_up).
bool init_radiance(Plan& p, std::string& err) {
  const int nz = p.nz, nxy = p.nx1 * p.ny1, nlayer = nz - 1, npart = p.npart;
  const float skyradalb = sky_flux_weight(p);
  p.radiance0.assign((size_t)p.npts * 4, 0.f);
  const LegTable lt = plan_leg_table(p);
  const int nchunk = 64;
  std::vector<char> bad(nchunk, 0);
  parallel_for(nchunk, [&](int c) {
    std::vector<float> od(nz), albs(nz), asym(nz), fl(3 * (size_t)(nlayer + 1));
    std::vector<double> work(4 * (size_t)(2 * nlayer + 3));
    double* wk = work.data();
    const size_t nw = 2 * (size_t)nlayer + 3;
    const int c0 = (int)((int64_t)nxy * c / nchunk), c1 = (int)((int64_t)nxy * (c + 1) / nchunk);
    for (int I = c0; I < c1; I++) {
      eddington_layers(lt, npart, p.d.deltam != 0, nz, p.zgrid.data(), p.total_ext.data(), p.extinct.data(),
                       p.albedo.data(), p.iphase.data(), p.phaseinterpwt.data(), (size_t)p.npts, I, od.data(),
                       albs.data(), asym.data());
      if (!eddrtf_column(nlayer, od.data(), albs.data(), asym.data(), p.d.solarflux, p.d.solarmu,
                         1.0f - p.d.gndalbedo, skyradalb, fl.data(), wk, wk + nw, wk + 2 * nw, wk + 3 * nw)) {
        bad[c] = 1;
        return;
      }
      column_radiance(nz, I, fl.data(), p.radiance0.data());
    }
  });
  for (char b : bad)
    if (b) {
      err = "EDDRTF: singular two-stream system";
      return false;
    }
  return true;
}

// ---------------------------------------------------------------------------
// Discrete ordinate trace geometry (BACK_INT_GRID3D_UNPOL shdomsub1.f:3736-3785).
// ---------------------------------------------------------------------------
void make_ordinates(Plan& p) {
  p.ords.clear();
  const double PI = (double)acosf(-1.0f);  // ACOS(-1.0) of a REAL argument
  const int c1 = p.gridptr[0 * 8 + 0], c8 = p.gridptr[0 * 8 + 7];
  const double eps = (double)(1.0e-3f * (p.gridpos[c8 * 3 + 2] - p.gridpos[c1 * 3 + 2]));
  for (int imu = 0; imu < p.nmu; imu++)
    for (int ip = 0; ip < p.nphi0[imu]; ip++) {
      Ordinate o;
      o.imu = imu;
      o.iphi = ip;
      o.mu = p.mu[imu];
      o.phi = p.phi[(size_t)imu * p.nphi0max + ip];
      o.wtdo = p.wtdo[(size_t)imu * p.nphi0max + ip];
      o.down = o.mu < 0.f;
      const double st = (double)sqrtf(1.0f - o.mu * o.mu);
      OrdGeom& g = o.g;
      g.cx = st * cos((double)o.phi + PI);
      g.cy = st * sin((double)o.phi + PI);
      g.cz = -(double)o.mu;
      if (fabs(g.cx) > (double)1.0e-5f) g.cxinv = 1.0 / g.cx; else { g.cx = 0.0; g.cxinv = 1.0e6; }
      if (fabs(g.cy) > (double)1.0e-5f) g.cyinv = 1.0 / g.cy; else { g.cy = 0.0; g.cyinv = 1.0e6; }
      g.czinv = 1.0 / g.cz;
      g.bitx = g.cx < 0.0 ? 1 : 0;
      g.bity = g.cy < 0.0 ? 1 : 0;
      g.bitz = g.cz < -(double)1.0e-3f ? 1 : 0;
      g.ioct = g.bitx + 2 * g.bity + 4 * g.bitz;
      g.eps = eps;
      p.ords.push_back(o);
    }
}

// ---------------------------------------------------------------------------
// Sweep order (SWEEPING_ORDER shdomsub1.f:3261-3351 with SWEEP_BASE_CELL
// :4529-4619 and, on a refined grid, SWEEP_NEXT_CELL :4624-4696) and the
// dependency levels that let all points of one level be integrated in
// parallel.
// ---------------------------------------------------------------------------
void sweeping_order(const Plan& p, int ioct0, std::vector<int32_t>& ord) {
  const int ioct = ioct0 + 1, nz = p.nz, nxc = p.nxc, nyc = p.nyc, bc = p.bcflag;
  int dix, six, eix, diy, siy, eiy, diz, siz, eiz;
  if (btest(ioct - 1, 0)) {
    dix = 1;
    if (btest(bc, 0)) { six = nxc; eix = std::max(nxc - 1, 1); } else { six = 1; eix = nxc; }
  } else {
    dix = -1;
    if (btest(bc, 0)) { six = 1; eix = std::min(2, nxc); } else { six = nxc; eix = 1; }
  }
  if (btest(ioct - 1, 1)) {
    diy = 1;
    if (btest(bc, 1)) { siy = nyc; eiy = std::max(nyc - 1, 1); } else { siy = 1; eiy = nyc; }
  } else {
    diy = -1;
    if (btest(bc, 1)) { siy = 1; eiy = std::min(2, nyc); } else { siy = nyc; eiy = 1; }
  }
  if (btest(ioct - 1, 2)) { diz = 1; siz = 1; eiz = nz - 1; } else { diz = -1; siz = nz - 1; eiz = 1; }
  std::vector<char> visited(p.npts, 0);
  ord.clear();
  ord.reserve(p.npts);
  const bool tree = !p.treeptr.empty();
  std::vector<int32_t> stack;
  int ix = six, iy = siy, iz = siz;
  for (;;) {
    // the end cells of this base cell's tree, depth first, the child on the
    // side the ray comes from first (SWEEP_NEXT_CELL :4624-4696)
    int cell = (iz + (nz - 1) * (iy - 1) + (nz - 1) * nyc * (ix - 1)) - 1;
    for (;;) {
      if (tree && p.treeptr[(size_t)cell * 2 + 1] >= 0) {
        const int idir = ((p.cellflags[cell] >> 2) & 3) - 1;
        const int c1 = p.treeptr[(size_t)cell * 2 + 1];
        if ((ioct0 >> idir) & 1) {
          stack.push_back(c1 + 1);
          cell = c1;
        } else {
          stack.push_back(c1);
          cell = c1 + 1;
        }
        continue;
      }
      for (int k = 0; k < 8; k++) {
        const int corner = k ^ (7 - ioct0);  // CORNDOG(k+1, IOCT) - 1
        const int ipt = p.gridptr[(size_t)cell * 8 + corner];
        if (visited[ipt]) continue;
        visited[ipt] = 1;
        ord.push_back((cell << 3) | corner);
      }
      if (stack.empty()) break;
      cell = stack.back();
      stack.pop_back();
    }
    if (ix == eix) {
      if (iy == eiy) {
        if (iz == eiz) break;
        iz += diz;
        iy = siy;
      } else {
        iy = (iy + diy + nyc - 1) % nyc + 1;
      }
      ix = six;
    } else {
      ix = (ix + dix + nxc - 1) % nxc + 1;
    }
  }
}

// One octant's sweep order, dependency levels and level lists.
static int build_octant(const Plan& p, int o, const std::vector<char>& top, const std::vector<char>& bot,
                        OctantPlan& op, std::string& err) {
  GridTopo g{p.gridpos.data(), p.gridptr.data(), p.neighptr.data(), p.cellflags.data(), p.npts, p.ncells};
  const bool tree = !p.treeptr.empty();
  if (tree) g.treeptr = p.treeptr.data();
  {
    sweeping_order(p, o, op.sweepord);
    if ((int)op.sweepord.size() != p.npts) {
      err = "SWEEPING_ORDER did not visit every grid point";
      return AT3D_ERR_NUMERIC;
    }
    op.rank.assign(p.npts, 0);
    for (int r = 0; r < p.npts; r++) op.rank[p.gridptr[(size_t)(op.sweepord[r] >> 3) * 8 + (op.sweepord[r] & 7)]] = r;
    const bool up = (o >> 2) & 1;  // trace-back ray goes -z: upwelling ordinate
    const std::vector<char>& preset = up ? bot : top;
    std::vector<const Ordinate*> ords;
    for (const Ordinate& od : p.ords)
      if (od.g.ioct == o) ords.push_back(&od);
    op.level.assign(p.npts, 0);
    op.simple_pt.assign(p.npts, 0);
    op.nslow = 0;
    const int bitx = o & 1, bity = (o >> 1) & 1, bitz = (o >> 2) & 1;
    for (int r = 0; r < p.npts; r++) {
      const int cell = op.sweepord[r] >> 3, corner = op.sweepord[r] & 7;
      const int pt = p.gridptr[(size_t)cell * 8 + corner];
      if (preset[pt]) continue;
      auto valid = [&](int q) { return preset[q] || op.rank[q] < r; };
      int lvl = -1;
      bool simple = false;
      if (corner == o) {
        // one-cell characteristic for every ordinate of the octant if all far
        // faces are known (and are the cell's own: a face onto smaller cells
        // takes its corners from the cell across)
        const int8_t fl = p.cellflags[cell];
        int faces[3], nf = 0;
        if (!(fl & 1)) faces[nf++] = 1 - bitx;
        if (!(fl & 2)) faces[nf++] = 3 - bity;
        faces[nf++] = 5 - bitz;
        bool ok = true;
        int mx = 0;
        for (int f = 0; f < nf && ok; f++) {
          if (tree && nb_smaller(p.neighptr[(size_t)cell * 6 + faces[f]])) {
            ok = false;
            break;
          }
          for (int c = 0; c < 4; c++) {
            const int q = p.gridptr[(size_t)cell * 8 + face_corner(faces[f], c)];
            if (!valid(q)) { ok = false; break; }
            mx = std::max(mx, op.level[q]);
          }
        }
        if (ok) {
          lvl = mx + 1;
          // a regular base cell: the GPU may use analytic geometry
          simple = (fl & 3) == 0 && (!tree || p.treeptr[(size_t)cell * 2] < 0);
        }
      }
      if (lvl < 0) {
        op.nslow++;
        int mx = 0;
        for (const Ordinate* od : ords) {
          int icell = cell;
          double xe = p.gridpos[pt * 3], ye = p.gridpos[pt * 3 + 1], ze = p.gridpos[pt * 3 + 2];
          bool done = false;
          for (int step = 0; step <= p.ncells && !done; step++) {
            TraceStep t;
            if (!trace_step(g, od->g, icell, xe, ye, ze, t)) {
              err = "BACK_INT_GRID: negative path length (SO<0)";
              return AT3D_ERR_NUMERIC;
            }
            const bool vf = valid(t.i[0]) && valid(t.i[1]) && valid(t.i[2]) && valid(t.i[3]);
            if (t.next < 0 || vf) {
              if (!vf) {
                err = "BACK_INT_GRID: reached a boundary face without valid radiances";
                return AT3D_ERR_NUMERIC;
              }
              for (int c = 0; c < 4; c++) mx = std::max(mx, op.level[t.i[c]]);
              done = true;
            } else {
              icell = t.next;
            }
          }
          if (!done) {
            err = "BACK_INT_GRID: characteristic did not terminate";
            return AT3D_ERR_NUMERIC;
          }
        }
        lvl = mx + 1;
      }
      op.level[pt] = lvl;
      if (simple) op.simple_pt[pt] = 1;
    }
    // level lists (level >= 1), ordered by (level, rank)
    int nl = 0;
    for (int i = 0; i < p.npts; i++) nl = std::max(nl, op.level[i]);
    op.nlevels = nl;
    std::vector<int32_t> cnt(nl + 2, 0);
    for (int r = 0; r < p.npts; r++) {
      const int pt = p.gridptr[(size_t)(op.sweepord[r] >> 3) * 8 + (op.sweepord[r] & 7)];
      if (op.level[pt] > 0) cnt[op.level[pt]]++;
    }
    op.lvl_off.assign(nl + 1, 0);  // lvl_off[L-1] .. lvl_off[L] for level L
    for (int L = 1; L <= nl; L++) op.lvl_off[L] = op.lvl_off[L - 1] + cnt[L];
    op.lvl_pts.assign(op.lvl_off[nl], 0);
    op.lvl_cell.assign(op.lvl_off[nl], 0);
    std::vector<int32_t> fill(op.lvl_off.begin(), op.lvl_off.end());
    for (int r = 0; r < p.npts; r++) {
      const int cell = op.sweepord[r] >> 3;
      const int pt = p.gridptr[(size_t)cell * 8 + (op.sweepord[r] & 7)];
      const int L = op.level[pt];
      if (L <= 0) continue;
      const int pos = fill[L - 1]++;
      // one-cell characteristics in regular cells are stored as the point index,
      // every other point as ~point (traced by the general path)
      op.lvl_pts[pos] = op.simple_pt[pt] ? pt : ~pt;
      op.lvl_cell[pos] = cell;
    }
  }
  return AT3D_OK;
}

// The sweep order and levels depend only on the grid and the ordinate
// directions, which a stream of solves on a few grids and angle sets repeats:
// the last few are kept, keyed by every input they are built from.
namespace {
std::mutex level_cache_mutex;
std::list<std::pair<std::vector<uint32_t>, std::shared_ptr<const std::array<OctantPlan, 8>>>> level_cache;
constexpr size_t LEVEL_CACHE_ENTRIES = 16;
constexpr size_t LEVEL_CACHE_BYTES = size_t(2) << 30;   // host memory the kept entries may hold

size_t octants_bytes(const std::array<OctantPlan, 8>& a) {
  size_t b = 0;
  for (const OctantPlan& o : a)
    b += 4 * (o.sweepord.size() + o.rank.size() + o.level.size() + o.lvl_pts.size() + o.lvl_cell.size() +
              o.lvl_off.size() + o.ring_level.size() + o.ring_off.size() + o.ring_code.size() +
              o.ring_cell.size() + o.ring_pos.size()) +
         o.simple_pt.size();
  return b;
}

std::vector<uint32_t> level_key(const Plan& p) {
  std::vector<uint32_t> k = {(uint32_t)p.npts, (uint32_t)p.ncells, (uint32_t)p.nx1, (uint32_t)p.ny1,
                             (uint32_t)p.nz, (uint32_t)p.bcflag, (uint32_t)p.ipflag, (uint32_t)p.nang};
  auto bits = [&](const std::vector<float>& v) {
    k.push_back((uint32_t)v.size());
    for (float f : v) {
      uint32_t u;
      memcpy(&u, &f, 4);
      k.push_back(u);
    }
  };
  bits(p.xgrid);
  bits(p.ygrid);
  bits(p.zgrid);
  for (const Ordinate& o : p.ords) {
    uint32_t u[3];
    memcpy(&u[0], &o.mu, 4);
    memcpy(&u[1], &o.phi, 4);
    u[2] = (uint32_t)o.g.ioct;
    k.insert(k.end(), u, u + 3);
  }
  return k;
}
}  // namespace

// BOUNDARY_PNTS (shdomsub1.f:2173-2215): the points on the top and bottom
// boundaries, in point order.
void boundary_points(Plan& p) {
  p.bctop.clear();
  p.bcbot.clear();
  for (int i = 0; i < p.npts; i++) {
    if (p.gridpos[i * 3 + 2] >= p.zgrid[p.nz - 1]) p.bctop.push_back(i);
    if (p.gridpos[i * 3 + 2] <= p.zgrid[0]) p.bcbot.push_back(i);
  }
}

int build_levels(Plan& p, std::string& err) {
  // a refined grid is its solve's own: its levels are neither looked up nor kept
  const bool tree = !p.treeptr.empty();
  const std::vector<uint32_t> key = tree ? std::vector<uint32_t>() : level_key(p);
  if (!tree) {
    std::lock_guard<std::mutex> l(level_cache_mutex);
    for (auto it = level_cache.begin(); it != level_cache.end(); ++it)
      if (it->first == key) {
        p.oct_store = it->second;
        level_cache.splice(level_cache.begin(), level_cache, it);
        break;
      }
  }
  if (!p.oct_store) {
    std::vector<char> top(p.npts, 0), bot(p.npts, 0);
    for (int i : p.bctop) top[i] = 1;
    for (int i : p.bcbot) bot[i] = 1;
    auto octs = std::make_shared<std::array<OctantPlan, 8>>();
    std::vector<std::string> errs(8);
    std::vector<int> rcs(8, AT3D_OK);
    bool used[8] = {};
    for (const Ordinate& od : p.ords) used[od.g.ioct] = true;
    parallel_for(8, [&](int o) {
      rcs[o] = build_octant(p, o, top, bot, (*octs)[o], errs[o]);
      // (the level-ordered sweep serves base grids only)
      if (rcs[o] == AT3D_OK && used[o] && !tree && !ring_octant_base(p, o, (*octs)[o], errs[o]))
        rcs[o] = AT3D_ERR_NUMERIC;
    });
    for (int o = 0; o < 8; o++)
      if (rcs[o]) {
        err = errs[o];
        return rcs[o];
      }
    p.oct_store = octs;
    if (tree) {
      p.oct = p.oct_store->data();
      p.nlevels_max = 0;
      for (int o = 0; o < 8; o++) p.nlevels_max = std::max(p.nlevels_max, p.oct[o].nlevels);
      return AT3D_OK;
    }
    std::lock_guard<std::mutex> l(level_cache_mutex);
    level_cache.emplace_front(key, p.oct_store);
    size_t total = 0;
    for (const auto& e : level_cache) total += octants_bytes(*e.second);
    while (level_cache.size() > 1 &&
           (level_cache.size() > LEVEL_CACHE_ENTRIES || total > LEVEL_CACHE_BYTES)) {
      total -= octants_bytes(*level_cache.back().second);
      level_cache.pop_back();
    }
  }
  p.oct = p.oct_store->data();
  p.nlevels_max = 0;
  for (int o = 0; o < 8; o++) p.nlevels_max = std::max(p.nlevels_max, p.oct[o].nlevels);
  return AT3D_OK;
}

// ---------------------------------------------------------------------------
// Rendering constants.  PRECOMPUTE_PHASE_CHECK (shdomsub4.f:2388-2492): the
// untruncated phase function of every table entry on NSCATANGLE equispaced
// scattering angles, with the reference's negativity check.  Per point, the
// TMS terms of COMPUTE_SOURCE_1CELL_UNPOL (shdomsub2.f:3096-3175): the
// truncated single-scatter weights sum_species DA * LEGENT_l / (1-f) and the
// (phase table, DA * w / (1-f)) pairs of the exact single-scatter term.
// ---------------------------------------------------------------------------
int render_point_constants(Plan& p);

int render_constants(Plan& p, std::string& err) {
  const int nsa = p.nscatangle, nlegp = p.nlegp, np = p.numphase;
  p.phasetab.assign((size_t)np * nsa, 0.f);
  const double PI = acos(-1.0);
  const double ofourpi = 1.0 / (4.0 * PI);
  std::vector<double> pl(nlegp + 1);
  for (int j = 0; j < nsa; j++) {
    const double x = cos(PI * (double)j / (double)(nsa - 1));
    // WIGNERFCT(X, NLEG, 0, 0): Legendre polynomials by recurrence
    pl[0] = 1.0;
    if (nlegp >= 1) pl[1] = x;
    for (int n = 1; n <= nlegp - 1; n++) {
      const double f1 = (double)(2 * n + 1) * x / (double)(n + 1);
      const double f2 = (double)n / (double)(n + 1);
      pl[n + 1] = f1 * pl[n] - f2 * pl[n - 1];
    }
    for (int iph = 0; iph < np; iph++) {
      double a1 = 0.0;
      for (int l = 0; l <= nlegp; l++) {
        const float unsc = p.legenp[(size_t)iph * (nlegp + 1) + l] / (float)(2 * l + 1);
        a1 += (2.0 * l + 1.0) * (double)unsc * pl[l];
      }
      if (a1 <= 0.0 && p.phasetab_err.empty()) {
        // reported when rendering, where the reference raises it
        char buf[200];
        snprintf(buf, sizeof buf,
                 "PRECOMPUTE_PHASE_CHECK: negative phase function for tabulated phase function: "
                 "IPH %d J %d A1 %.10g", iph + 1, j + 1, a1);
        p.phasetab_err = buf;
      }
      p.phasetab[(size_t)iph * nsa + j] = (float)(a1 * ofourpi);
    }
  }
  if (p.nst == 3) {
    // the I-Q element: -sum (2l+1) b1_l d^l_{20}(cos T) / (4 pi) (PHASETAB(2))
    p.phasetab2.assign((size_t)np * nsa, 0.f);
    std::vector<double> d2;
    for (int j = 0; j < nsa; j++) {
      const double x = cos(PI * (double)j / (double)(nsa - 1));
      wigner_d(x, nlegp, 2, 0, d2);
      for (int iph = 0; iph < np; iph++) {
        double b1 = 0.0;
        for (int l = 0; l <= nlegp; l++) {
          const float unsc = p.legenp6[((size_t)4 * np + iph) * (nlegp + 1) + l] / (float)(2 * l + 1);
          b1 = b1 - (2.0 * l + 1.0) * (double)unsc * d2[l];
        }
        p.phasetab2[(size_t)iph * nsa + j] = (float)(b1 * ofourpi);
      }
    }
  }
  if (!p.host_medium) return AT3D_OK;
  return render_point_constants(p);
}

// The per-point single-scatter constants (TMS terms and the exact
// single-scatter pairs, CSR in point order) of every grid point.
int render_point_constants(Plan& p) {
  const int ML = p.ml, npts = p.npts, npart = p.npart;
  const float secmu0 = (float)(1.0 / fabs((double)p.d.solarmu));
  p.tms.assign((size_t)npts * (ML + 1), 0.f);
  p.ss_off.assign(npts + 1, 0);
  p.ss_iph.clear();
  p.ss_coef.clear();
  if (!p.d.deltam) return AT3D_OK;
  const LegTable lt = plan_leg_table(p);
  // per point: the TMS terms and its (phase, coefficient) list; points in
  // chunks over the host threads, each chunk's lists joined in point order
  const int nchunk = 64;
  std::vector<std::vector<int32_t>> c_iph(nchunk);
  std::vector<std::vector<float>> c_coef(nchunk);
  std::vector<int32_t> cnt(npts, 0);
  parallel_for(nchunk, [&](int c) {
    std::vector<int32_t> ib(8 * (size_t)npart);
    std::vector<float> cb(8 * (size_t)npart);
    const int i0 = (int)((int64_t)npts * c / nchunk), i1 = (int)((int64_t)npts * (c + 1) / nchunk);
    for (int ip = i0; ip < i1; ip++) {
      const int n = point_render_constants(lt, npart, p.total_ext[ip], &p.extinct[ip], &p.albedo[ip],
                                           &p.iphase[(size_t)ip * p.npw], &p.phaseinterpwt[(size_t)ip * p.npw],
                                           (size_t)npts, secmu0, &p.tms[(size_t)ip * (ML + 1)], ib.data(),
                                           cb.data());
      c_iph[c].insert(c_iph[c].end(), ib.begin(), ib.begin() + n);
      c_coef[c].insert(c_coef[c].end(), cb.begin(), cb.begin() + n);
      cnt[ip] = n;
    }
  });
  for (int ip = 0; ip < npts; ip++) p.ss_off[ip + 1] = p.ss_off[ip] + cnt[ip];
  for (int c = 0; c < nchunk; c++) {
    p.ss_iph.insert(p.ss_iph.end(), c_iph[c].begin(), c_iph[c].end());
    p.ss_coef.insert(p.ss_coef.end(), c_coef[c].begin(), c_coef[c].end());
  }
  return AT3D_OK;
}

}  // namespace

void ylm_direction(float mu, float phi, int ml, int mm, float* yr) { ylm_unpol(mu, phi, ml, mm, yr); }

// Wigner d^n_{m,m1}(x): the value at n0 = max(m, m1) (DMM1_N0) and the
// three-term recurrence in n above it.
void wigner_d(double x, int nrank, int m, int m1, std::vector<double>& d) {
  d.assign(nrank + 1, 0.0);
  const int n0 = std::max(m, m1);
  if (n0 > nrank) return;
  if (n0 == 0) {
    d[0] = 1.0;
    if (nrank >= 1) d[1] = x;
    for (int n = 1; n <= nrank - 1; n++) {
      const double f1 = (double)(2 * n + 1) * x / (double)(n + 1);
      const double f2 = (double)n / (double)(n + 1);
      d[n + 1] = f1 * d[n] - f2 * d[n - 1];
    }
    return;
  }
  d[n0] = dmm1_n0(x, m, m1);
  for (int n = n0; n <= nrank - 1; n++) {
    const double rm = sqrt((double)((n + 1) * (n + 1) - m * m));
    const double rm1 = sqrt((double)((n + 1) * (n + 1) - m1 * m1));
    double f1 = (double)(n * (n + 1)) * x - (double)(m * m1);
    f1 = f1 / rm / rm1 * (double)(2 * n + 1) / (double)n;
    double f2 = sqrt((double)(n * n - m * m)) * sqrt((double)(n * n - m1 * m1));
    f2 = f2 / rm / rm1 * (double)(n + 1) / (double)n;
    d[n + 1] = f1 * d[n] - f2 * d[n - 1];
  }
}

// YLMALL with NSTLEG = 6 (shdomsub2.f:4244-4359): the zenith functions of
// plm_pol times cos / sin of m phi in the combinations the polarized basis
// uses, [nlm][6].
void ylm_direction_pol(float mu, float phi, int ml, int mm, float* yr6) {
  const double fct = 1.0 / sqrt(2.0 * acos(-1.0));
  std::vector<double> dm0, dp, dm;
  for (int m = 0; m <= mm; m++) {
    wigner_dm0_plm((double)mu, ml, m, dm0);
    wigner_d2pm((double)mu, ml, m, dp, dm);
    double c = 1.0, sn = 0.0;
    if (m > 0) {
      c = (double)cosf((float)m * phi);
      sn = (double)sinf((float)m * phi);
    }
    for (int l = m; l <= ml; l++) {
      const double p1 = fct * dm0[l];
      const double p2 = -0.5 * fct * (dp[l] + dm[l]);
      const double p3 = -0.5 * fct * (dp[l] - dm[l]);
      float* a = yr6 + (size_t)sh_index(l, m, mm) * 6;
      if (m == 0) {
        a[0] = (float)p1; a[1] = (float)p2; a[2] = (float)p2;
        a[3] = (float)p1; a[4] = (float)p3; a[5] = (float)p3;
        continue;
      }
      a[0] = (float)(p1 * c - p1 * sn); a[1] = (float)(p2 * c - p2 * sn); a[2] = (float)(p2 * c + p2 * sn);
      a[3] = (float)(p1 * c + p1 * sn); a[4] = (float)(p3 * c - p3 * sn); a[5] = (float)(p3 * c + p3 * sn);
      float* b = yr6 + (size_t)sh_index(l, -m, mm) * 6;
      b[0] = (float)(p1 * sn + p1 * c); b[1] = (float)(p2 * sn + p2 * c); b[2] = (float)(p2 * sn - p2 * c);
      b[3] = (float)(p1 * sn - p1 * c); b[4] = (float)(p3 * sn + p3 * c); b[5] = (float)(p3 * sn - p3 * c);
    }
  }
}

// COMPUTE_SOURCE (shdomsub1.f:967-1611) forms source term J as albedo x
// chi_l(J) x (...), so a point whose albedo-weighted phase terms vanish above
// order l has no source terms above l.
void plan_source_class(const Plan& p, std::vector<int32_t>& cls) {
  const int npts = p.npts;
  cls.assign(npts, 0);
  if (p.chi.size() < (size_t)npts * (p.ml + 1) || p.alb.size() < (size_t)npts) return;
  for (int i = 0; i < npts; i++)
    for (int l = p.ml; l > 0; l--)
      if (p.alb[i] * p.chi[(size_t)i * (p.ml + 1) + l] != 0.f) {
        cls[i] = l;
        break;
      }
}

void plan_mixed_legen(const Plan& p, int ipa, int ip, std::vector<float>& out) {
  mixed_legen(p, ipa, ip, out);
}

int plan_direct_host(const Plan& p, std::vector<float>& dirflux, std::string& err) {
  std::vector<float> extdirp;
  DirectBeamGeom beam;
  direct_beam_setup(p, extdirp, beam);
  std::vector<float> df(p.npts, 0.f);
  const int nchunk = 64;
  std::vector<char> bad(nchunk, 0);
  parallel_for(nchunk, [&](int c) {
    const int i0 = (int)((int64_t)p.npts * c / nchunk), i1 = (int)((int64_t)p.npts * (c + 1) / nchunk);
    for (int ip = i0; ip < i1; ip++)
      if (!direct_beam_point(beam, p.gridpos[ip * 3], p.gridpos[ip * 3 + 1], p.gridpos[ip * 3 + 2], df[ip],
                             nullptr))
        bad[c] = 1;
  });
  for (char b : bad)
    if (b) {
      err = "DIRECT_BEAM_PROP: ray left the property grid";
      return AT3D_ERR_NUMERIC;
    }
  dirflux.swap(df);
  return AT3D_OK;
}

int plan_sun_paths(const Plan& p, std::vector<int32_t>& off, std::vector<int32_t>& ptr,
                   std::vector<float>& val, std::string& err) {
  std::vector<float> extdirp;
  DirectBeamGeom beam;
  direct_beam_setup(p, extdirp, beam);
  off.assign(p.npts + 1, 0);
  ptr.clear();
  val.clear();
  std::vector<std::pair<int32_t, float>> rec;
  for (int ip = 0; ip < p.npts; ip++) {
    off[ip] = (int32_t)ptr.size();
    rec.clear();
    float df;
    if (!direct_beam_point(beam, p.gridpos[ip * 3], p.gridpos[ip * 3 + 1], p.gridpos[ip * 3 + 2], df, &rec)) {
      err = "DIRECT_BEAM_AND_PATHS_PROP: ray left the property grid";
      return AT3D_ERR_NUMERIC;
    }
    for (auto& e : rec) {
      ptr.push_back(e.first);
      val.push_back(e.second);
    }
  }
  off[p.npts] = (int32_t)ptr.size();
  return AT3D_OK;
}

int build_plan(const at3d_scene_desc& desc, Plan& p, std::string& err) {
  p.d = desc;
  const at3d_scene_desc& d = desc;
  if (d.npx < 2 || d.npy < 2 || d.npz < 2) {
    err = "grid must be at least 2x2x2 (1D/2D independent-pixel grids are not supported)";
    return AT3D_ERR_ARG;
  }
  if (d.transmin != 1.0f) {
    err = "transmin must be 1.0 (one-cell short characteristics)";
    return AT3D_ERR_ARG;
  }
  if (d.ip_flag != 0) {
    err = "independent pixel modes are not supported (ip_flag must be 0)";
    return AT3D_ERR_ARG;
  }
  if (!(d.solarmu < 0.0f && d.solarmu >= -1.0f)) {
    err = "solarmu must be in [-1, 0)";
    return AT3D_ERR_ARG;
  }
  if (d.npart < 1 || d.numphase < 1 || d.nlegp < 0) {
    err = "need at least one species and one phase function";
    return AT3D_ERR_ARG;
  }
  const int npg = d.npx * d.npy * d.npz;
  // phase tables per property point (num_micro, MAXNMICRO; 0 is taken as 1)
  p.nmicro = d.nmicro == 0 ? 1 : d.nmicro;
  if (p.nmicro < 1 || p.nmicro > MAX_NMICRO) {
    err = "nmicro must be 1 .. " + std::to_string(MAX_NMICRO);
    return AT3D_ERR_ARG;
  }
  if (p.nmicro > 1 && d.splitacc > 0.0f) {
    err = "more than one phase table per property point runs on the base grid: split_accuracy must be 0";
    return AT3D_ERR_ARG;
  }
  p.npw = 8 * p.nmicro;
  if (d.interpmethod != 0 && d.interpmethod != 1) {
    err = "interpmethod must be 0 ('ON') or 1 ('OO')";
    return AT3D_ERR_ARG;
  }
  p.interp_o = d.interpmethod == 1;
  if (p.interp_o && p.nmicro != 1) {
    err = "INTERPMETHOD 'OO' takes one phase table per property point (nmicro = 1)";
    return AT3D_ERR_ARG;
  }
  p.zlevels.assign(d.zlevels, d.zlevels + d.npz);
  for (int i = 1; i < d.npz; i++)
    if (!(p.zlevels[i] > p.zlevels[i - 1])) {
      err = "zlevels must increase";
      return AT3D_ERR_ARG;
    }
  p.npart = d.npart;
  p.extinctp.assign(d.extinctp, d.extinctp + (size_t)npg * d.npart);
  p.albedop.assign(d.albedop, d.albedop + (size_t)npg * d.npart);
  p.iphasep.assign(d.iphasep, d.iphasep + (size_t)npg * d.npart * p.nmicro);
  p.phasewtp.assign(d.phasewtp, d.phasewtp + (size_t)npg * d.npart * p.nmicro);
  for (int v : p.iphasep)
    if (v < 1 || v > d.numphase) {
      err = "Phase function indices are out of bounds.";
      return AT3D_ERR_ARG;
    }
  // solver.py:1970-1990 and 2190-2200
  p.nmu = std::max(2, 2 * ((d.nmu + 1) / 2));
  p.nphi = std::max(1, d.nphi);
  p.ml = p.nmu - 1;
  p.mm = std::max(0, p.nphi / 2 - 1);
  p.nlm = (2 * p.mm + 1) * (p.ml + 1) - p.mm * (p.mm + 1);
  if (p.nlm < 4) {
    err = "Insufficient Spherical Harmonics (NLM < 4)";
    return AT3D_ERR_ARG;
  }
  p.nphi0max = p.nphi;
  p.nleg = d.deltam ? p.ml + 1 : p.ml;
  p.numphase = d.numphase;
  // pad the phase table to nleg (solver.py:2385-2395)
  const int nlegp_in = d.nlegp;
  p.nlegp = std::max(nlegp_in, p.nleg);
  p.nscatangle = std::max(36, std::min(721, 2 * p.nlegp));
  // Stokes components: the polarized input table holds six elements per
  // order; element 1 (a1) is the scalar table
  p.nst = d.nstokes == 0 ? 1 : d.nstokes;
  if (p.nst != 1 && p.nst != 3) {
    err = "nstokes must be 1 or 3";
    return AT3D_ERR_ARG;
  }
  const int nstleg = p.nst == 3 ? 6 : 1;
  p.legenp.assign((size_t)p.numphase * (p.nlegp + 1), 0.f);
  for (int iph = 0; iph < p.numphase; iph++)
    for (int l = 0; l <= nlegp_in; l++)
      p.legenp[(size_t)iph * (p.nlegp + 1) + l] = d.legenp[((size_t)iph * (nlegp_in + 1) + l) * nstleg];
  if (p.nst == 3) {
    p.legenp6.assign((size_t)6 * p.numphase * (p.nlegp + 1), 0.f);
    for (int e = 0; e < 6; e++)
      for (int iph = 0; iph < p.numphase; iph++)
        for (int l = 0; l <= nlegp_in; l++)
          p.legenp6[((size_t)e * p.numphase + iph) * (p.nlegp + 1) + l] =
              d.legenp[((size_t)iph * (nlegp_in + 1) + l) * 6 + e];
  }
  // _setup_grid (solver.py:2036-2168)
  p.nx = d.npx;
  p.ny = d.npy;
  p.nz = d.npz;
  p.ipflag = d.ip_flag;
  p.bcflag = d.open_bc ? 3 : 0;
  p.nx1 = p.nx + 1;
  p.ny1 = p.ny + 1;
  if ((p.bcflag & 5) || (p.ipflag & 1)) p.nx1 -= 1;
  if ((p.bcflag & 7) || (p.ipflag & 2)) p.ny1 -= 1;
  make_grid(p);
  int rc = make_angle_set(p, err);
  if (rc) return rc;
  make_sh_tables(p);
  // the polarized per-point terms are made on the host
  if (d.device_medium && p.nst == 1) {
    // the state makes the per-point medium on the device
    medium_scalars(p);
    prepare_legen(p);
    p.medium_on_device = true;
    p.host_medium = false;
  } else {
    interp_optics(p);
    source_constants(p);
    // MAKE_DIRECT: the direct flux is made per solve on the device (or on the
    // host by plan_direct_host when asked for); the plan keeps the constants
    // that scale with it per unit direct flux
    if (!init_radiance(p, err)) return AT3D_ERR_NUMERIC;
    if (p.nst == 3) polarized_constants(p);
  }
  boundary_points(p);
  make_ordinates(p);
  rc = build_levels(p, err);
  if (rc) return rc;
  rc = render_constants(p, err);
  if (rc) return rc;
  // the memory parameters bound the adaptive grid (0: the base grid's sizes)
  p.nbpts = p.npts;
  p.nbcells = p.ncells;
  p.maxig = std::max(d.maxig, p.npts);
  p.maxic = std::max(d.maxic, p.ncells);
  p.maxiv = d.maxiv > 0 ? d.maxiv : p.nlm * p.npts;
  p.maxido = d.maxido > 0 ? d.maxido : p.maxig * p.nphi0max;
  return AT3D_OK;
}

int plan_set_grid(Plan& p, int npts, int ncells, const float* gridpos, const int32_t* gridptr,
                  const int32_t* neighptr, const int32_t* treeptr, const int8_t* cellflags, std::string& err) {
  // the grid must be a tree of end cells over the base cells whose pointers
  // stay in range, and descents must end (children come after their parent)
  if (npts < p.nbpts || ncells < p.nbcells || npts > p.maxig || ncells > p.maxic) {
    err = "Cannot load solution as loaded grid has more points or cells than supported by the current "
          "RTE's memory parameters (or fewer than its base grid)";
    return AT3D_ERR_ARG;
  }
  for (size_t i = 0; i < (size_t)ncells * 8; i++)
    if (gridptr[i] < 0 || gridptr[i] >= npts) {
      err = "loaded grid: corner pointer out of range";
      return AT3D_ERR_ARG;
    }
  for (size_t i = 0; i < (size_t)ncells * 6; i++)
    if (nb_cell(neighptr[i]) >= ncells || neighptr[i] < -ncells - 1) {
      err = "loaded grid: neighbour pointer out of range";
      return AT3D_ERR_ARG;
    }
  for (int c = 0; c < ncells; c++) {
    const int par = treeptr[(size_t)c * 2], ch = treeptr[(size_t)c * 2 + 1];
    const int dir = (cellflags[c] >> 2) & 3;
    const bool bad_parent = par >= c || par < -1 || (c < p.nbcells) != (par < 0);
    const bool bad_child = ch >= 0 && (ch <= c || ch + 1 >= ncells || dir < 1 || treeptr[(size_t)ch * 2] != c ||
                                       treeptr[(size_t)(ch + 1) * 2] != c);
    if (bad_parent || bad_child || ch < -1) {
      err = "loaded grid: inconsistent cell tree";
      return AT3D_ERR_ARG;
    }
  }
  p.npts = npts;
  p.ncells = ncells;
  p.gridpos.assign(gridpos, gridpos + (size_t)npts * 3);
  p.gridptr.assign(gridptr, gridptr + (size_t)ncells * 8);
  p.neighptr.assign(neighptr, neighptr + (size_t)ncells * 6);
  p.treeptr.assign(treeptr, treeptr + (size_t)ncells * 2);
  p.cellflags.assign(cellflags, cellflags + ncells);
  // the optical properties of every point from the property grid
  // (_prepare_optical_properties after load_solution, solver.py:1595)
  interp_optics(p);
  return plan_refresh_points(p, err);
}

int ensure_host_medium(const Plan& pc, std::string& err) {
  static std::mutex m;
  std::lock_guard<std::mutex> lock(m);
  if (pc.host_medium) return AT3D_OK;
  Plan& p = const_cast<Plan&>(pc);   // a plan object is never const itself
  interp_optics(p);
  source_constants(p);
  if (!init_radiance(p, err)) return AT3D_ERR_NUMERIC;
  render_point_constants(p);
  p.host_medium = true;
  return AT3D_OK;
}

int plan_refresh_points(Plan& p, std::string& err) {
  source_constants(p);
  if (p.nst == 3) polarized_constants(p);
  // the two-stream start is a base-grid quantity; new points start from zero
  p.radiance0.resize((size_t)p.npts * 4, 0.f);
  boundary_points(p);
  p.oct_store.reset();
  p.oct = nullptr;
  int rc = build_levels(p, err);
  if (rc) return rc;
  return render_constants(p, err);
}

}  // namespace at3d
