// Host-side preparation of a Levis-approximation gradient: what
// RTE.calculate_microphysical_partial_derivatives and
// RTE.calculate_direct_beam_derivative (at3d/solver.py:1266-1516) compute
// before LEVISAPPROX_GRADIENT runs.
#include "grad_plan.h"

#include <math.h>

#include <algorithm>

namespace at3d {

namespace {

// COMPUTE_INTERP_WEIGHTS (shdomsub4.f:3082-3169): trilinear weights and
// 0-based pointers of the 8 property-grid points around one RTE point.
bool interp_weights(const Plan& p, float X, float Y, float Z, float* wt, int32_t* ptr) {
  const at3d_scene_desc& d = p.d;
  const int npx = d.npx, npy = d.npy, npz = d.npz;
  const float* ZL = p.zlevels.data();
  int il = 0, iu = npz;
  while (iu - il > 1) {
    const int im = (iu + il) / 2;
    if (Z >= ZL[im - 1]) il = im; else iu = im;
  }
  const int iz = std::max(il, 1);
  double w = (double)(Z - ZL[iz - 1]) / (double)(ZL[iz] - ZL[iz - 1]);
  w = std::max(std::min(w, 1.0), 0.0);
  int ix = (int)((X - 0.0f) / d.delx) + 1;
  if (fabsf(X - 0.0f - (float)npx * d.delx) < 0.01f * d.delx) ix = npx;
  if (ix < 1 || ix > npx) return false;
  const int ixp = ix % npx + 1;
  double u = (double)(X - 0.0f - d.delx * (float)(ix - 1)) / (double)d.delx;
  u = std::max(std::min(u, 1.0), 0.0);
  if (u < 1.0e-5) u = 0.0;
  if (u > 1.0 - 1.0e-5) u = 1.0;
  int iy = (int)((Y - 0.0f) / d.dely) + 1;
  if (fabsf(Y - 0.0f - (float)npy * d.dely) < 0.01f * d.dely) iy = npy;
  if (iy < 1 || iy > npy) return false;
  const int iyp = iy % npy + 1;
  double v = (double)(Y - 0.0f - d.dely * (float)(iy - 1)) / (double)d.dely;
  v = std::max(std::min(v, 1.0), 0.0);
  if (v < 1.0e-5) v = 0.0;
  if (v > 1.0 - 1.0e-5) v = 1.0;
  wt[0] = (float)((1 - u) * (1 - v) * (1 - w));
  wt[1] = (float)(u * (1 - v) * (1 - w));
  wt[2] = (float)((1 - u) * v * (1 - w));
  wt[3] = (float)(u * v * (1 - w));
  wt[4] = (float)((1 - u) * (1 - v) * w);
  wt[5] = (float)(u * (1 - v) * w);
  wt[6] = (float)((1 - u) * v * w);
  wt[7] = (float)(u * v * w);
  ptr[0] = iz + npz * (iy - 1) + npz * npy * (ix - 1) - 1;
  ptr[1] = iz + npz * (iy - 1) + npz * npy * (ixp - 1) - 1;
  ptr[2] = iz + npz * (iyp - 1) + npz * npy * (ix - 1) - 1;
  ptr[3] = iz + npz * (iyp - 1) + npz * npy * (ixp - 1) - 1;
  for (int q = 0; q < 4; q++) ptr[q + 4] = ptr[q] + 1;
  return true;
}

}  // namespace

int build_grad_plan(const Plan& p, const at3d_deriv_desc& dd, GradPlan& g, std::string& err) {
  const at3d_scene_desc& d = p.d;
  const int npg = d.npx * d.npy * d.npz, npts = p.npts, ML = p.ml, nl = p.nleg + 1;
  const bool deltam = d.deltam != 0;
  if (dd.numder < 1) {
    err = "numder must be >= 1";
    return AT3D_ERR_ARG;
  }
  if (dd.dnumphase < 1) {
    err = "dnumphase must be >= 1";
    return AT3D_ERR_ARG;
  }
  g.numder = dd.numder;
  g.maxpg = npg;
  g.dnumphase = dd.dnumphase;
  g.partder.resize(g.numder);
  g.doexact.resize(g.numder);
  for (int i = 0; i < g.numder; i++) {
    if (dd.partder[i] < 1 || dd.partder[i] > p.npart) {
      err = "partder out of range";
      return AT3D_ERR_ARG;
    }
    g.partder[i] = dd.partder[i] - 1;
    g.doexact[i] = dd.doexact ? dd.doexact[i] : 1;
  }
  const size_t nd = (size_t)g.numder * npg;
  g.dext.assign(dd.dext, dd.dext + nd);
  g.dalb.assign(dd.dalb, dd.dalb + nd);
  g.dphasewtp.assign(nd, 0.f);
  if (dd.dphasewtp) g.dphasewtp.assign(dd.dphasewtp, dd.dphasewtp + nd);
  g.diphasep.assign(nd, 0);
  for (size_t i = 0; i < nd; i++) {
    int v = dd.diphasep ? dd.diphasep[i] : 1;
    if (v == 0) v = 1;   // points not covered by a derivative table (solver.py:1420-1425)
    if (v < 1 || v > g.dnumphase) {
      err = "derivative phase function indices are out of bounds";
      return AT3D_ERR_ARG;
    }
    g.diphasep[i] = v - 1;
  }
  // derivative Legendre table: padded to the solver's table length, l = 0 term
  // zeroed and divided by (2l+1) (solver.py:1427-1463)
  const int nlin = d.nlegp, nlp = p.nlegp;
  std::vector<float> dfull((size_t)g.dnumphase * (nlp + 1), 0.f);
  for (int t = 0; t < g.dnumphase; t++)
    for (int l = 1; l <= std::min(nlin, nlp); l++)
      dfull[(size_t)t * (nlp + 1) + l] =
          (dd.dleg ? dd.dleg[(size_t)t * (nlin + 1) + l] : 0.f) / (float)(2 * l + 1);
  g.dleg.assign((size_t)g.dnumphase * nl, 0.f);
  for (int t = 0; t < g.dnumphase; t++)
    for (int l = 0; l < nl && l <= nlp; l++) g.dleg[(size_t)t * nl + l] = dfull[(size_t)t * (nlp + 1) + l];
  // PRECOMPUTE_PHASE_CHECK_GRAD without the negativity check
  const int nsa = p.nscatangle;
  g.dphasetab.assign((size_t)g.dnumphase * nsa, 0.f);
  {
    const double PI = acos(-1.0), ofourpi = 1.0 / (4.0 * PI);
    std::vector<double> pl(nlp + 1);
    for (int j = 0; j < nsa; j++) {
      const double x = cos(PI * (double)j / (double)(nsa - 1));
      pl[0] = 1.0;
      if (nlp >= 1) pl[1] = x;
      for (int n = 1; n <= nlp - 1; n++)
        pl[n + 1] = (double)(2 * n + 1) * x / (double)(n + 1) * pl[n] - (double)n / (double)(n + 1) * pl[n - 1];
      for (int t = 0; t < g.dnumphase; t++) {
        double a1 = 0.0;
        for (int l = 0; l <= nlp; l++) a1 += (2.0 * l + 1.0) * (double)dfull[(size_t)t * (nlp + 1) + l] * pl[l];
        g.dphasetab[(size_t)t * nsa + j] = (float)(a1 * ofourpi);
      }
    }
  }
  // phase 1: interpolation weights (COMPUTE_INTERP_WEIGHTS)
  g.optinterpwt.assign((size_t)npts * 8, 0.f);
  g.interpptr.assign((size_t)npts * 8, 0);
  for (int ip = 0; ip < npts; ip++)
    if (!interp_weights(p, p.gridpos[ip * 3], p.gridpos[ip * 3 + 1], p.gridpos[ip * 3 + 2],
                        &g.optinterpwt[(size_t)ip * 8], &g.interpptr[(size_t)ip * 8])) {
      err = "TRILIN: Beyond X or Y domain";
      return AT3D_ERR_ARG;
    }
  // phase 2: FP, DFP and DEXTM on the property grid
  auto leg = [&](int iph0, int l) { return p.legen[(size_t)iph0 * nl + l]; };
  std::vector<float> fp(nd, 0.f), dfp(nd, 0.f);
  g.dextm.assign(nd, 0.f);
  for (int idr = 0; idr < g.numder; idr++) {
    const int ipa = g.partder[idr];
    for (int ib = 0; ib < npg; ib++) {
      const size_t pi = (size_t)ipa * npg + ib, di = (size_t)idr * npg + ib;
      float fpv = 0.f, dfpv = 0.f;
      if (deltam) {
        fpv = fpv + p.phasewtp[pi] * leg(p.iphasep[pi] - 1, ML + 1);
        if (g.doexact[idr] == 1)
          dfpv = dfpv + p.phasewtp[pi] * g.dleg[(size_t)g.diphasep[di] * nl + ML + 1];
        else
          dfpv = dfpv + g.dphasewtp[di] * leg(p.iphasep[pi] - 1, ML + 1);
      }
      fp[di] = fpv;
      dfp[di] = dfpv;
      const float ep = p.extinctp[pi], ap = p.albedop[pi];
      g.dextm[di] = g.dext[di] * (1 - fpv * ap) - g.dalb[di] * fpv * ep - ep * ap * dfpv;
    }
  }
  // phase 3: DALBM and DFJ on the RTE grid
  g.dalbm.assign((size_t)g.numder * npts * 8, 0.f);
  g.dfj.assign((size_t)g.numder * npts * 8, 0.f);
  for (int idr = 0; idr < g.numder; idr++) {
    const int ipa = g.partder[idr];
    for (int ip = 0; ip < npts; ip++) {
      float F = 0.f, albj;
      const size_t base = ((size_t)ipa * npts + ip) * 8;
      const float alb = p.albedo[(size_t)ipa * npts + ip];
      if (deltam) {
        if (p.phaseinterpwt[base] >= 0.999f) {
          F = leg(p.iphase[base] - 1, ML + 1);
        } else {
          for (int q = 0; q < 8; q++) {
            if (p.phaseinterpwt[base + q] < 1e-7f) continue;
            F = F + p.phaseinterpwt[base + q] * leg(p.iphase[base + q] - 1, ML + 1);
          }
        }
        albj = alb / (F * (alb - 1) + 1);
      } else {
        albj = alb;
      }
      const float divide = 1.0f / (1.0f - albj * F);
      for (int nb = 0; nb < 8; nb++) {
        const int ib = g.interpptr[(size_t)ip * 8 + nb];
        const size_t pi = (size_t)ipa * npg + ib, di = (size_t)idr * npg + ib;
        const float ep = p.extinctp[pi], ap = p.albedop[pi];
        const size_t o = ((size_t)idr * npts + ip) * 8 + nb;
        g.dalbm[o] = divide * (g.dext[di] * ((1 - F) * (ap - albj) + (albj - 1) * ap * (fp[di] - F)) +
                               g.dalb[di] * ((1 - F) * ep + (albj - 1) * ep * (fp[di] - F)) +
                               dfp[di] * (albj - 1) * ep * ap);
        g.dfj[o] = (g.dext[di] * (fp[di] - F) * ap + g.dalb[di] * (fp[di] - F) * ep + dfp[di] * ep * ap) / (1 - F);
      }
    }
  }
  // sun paths, transposed to the property grid.  The reference stores each
  // point's path in a column of LONGEST_PATH_PTS slots (the longest path) and
  // COMPUTE_DIRECT_BEAM_DERIV_ADJOINT (shdomsub4.f:4117-4143) reads a column
  // until its first empty slot, with no length bound: a point whose path fills
  // its column also takes the entries of the next point's column (and on, while
  // those are full too).  The effective path of a point reproduces that.
  std::vector<int32_t> off, ptr;
  std::vector<float> val;
  int rc = plan_sun_paths(p, off, ptr, val, err);
  if (rc) return rc;
  int32_t longest = 1;
  for (int ip = 0; ip < npts; ip++) longest = std::max(longest, off[ip + 1] - off[ip]);
  auto for_path = [&](int ip, auto&& f) {
    for (int q = ip; q < npts; q++) {
      for (int32_t e = off[q]; e < off[q + 1]; e++) f(ptr[e], val[e]);
      if (off[q + 1] - off[q] < longest) break;
    }
  };
  g.bp_off.assign(npg + 1, 0);
  for (int ip = 0; ip < npts; ip++) for_path(ip, [&](int32_t ib, float) { g.bp_off[ib + 1]++; });
  for (int i = 0; i < npg; i++) g.bp_off[i + 1] += g.bp_off[i];
  g.bp_ip.assign(g.bp_off[npg], 0);
  g.bp_w.assign(g.bp_off[npg], 0.f);
  std::vector<int32_t> fill(g.bp_off.begin(), g.bp_off.end() - 1);
  for (int ip = 0; ip < npts; ip++)
    for_path(ip, [&](int32_t ib, float w) {
      g.bp_ip[fill[ib]] = ip;
      g.bp_w[fill[ib]] = w;
      fill[ib]++;
    });
  return AT3D_OK;
}

}  // namespace at3d
