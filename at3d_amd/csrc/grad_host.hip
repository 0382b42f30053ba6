// Host orchestration of the Levis gradient (LEVISAPPROX_GRADIENT,
// shdomsub4.f:288-809, as driven by
// LevisApproxGradient.levis_approximation_grad, at3d/gradient.py:164-415):
// the adjoint branch for the gradient and cost, and the Jacobian branch's
// per-pixel derivatives at requested property points (:378-467) from the
// same ray traversal.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "grad_plan.h"
#include "host_plan.h"
#include "kernels.h"
#include "solver_state.h"

namespace at3d {

#define HIPCHK(x)                                                              \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      err = std::string(#x) + ": " + hipGetErrorString(e_);                    \
      return AT3D_ERR_DEVICE;                                                  \
    }                                                                          \
  } while (0)

namespace {

size_t rup(size_t v, size_t m) { return (v + m - 1) / m * m; }

// Per species and RTE point: the mixed phase function of
// COMPUTE_SOURCE_GRAD_1CELL (shdomsub4.f:1703-1723 for one species,
// :1823-1887 from the property grid for several), delta-M scaled (legt_s) and
// unscaled (legt_u), and the SINGSCATJ terms (:1805-1821) as (phase table,
// coefficient) pairs.
struct SpeciesTables {
  std::vector<float> legt_s, legt_u, f_s;
  std::vector<int8_t> scatj_ok;
  std::vector<int32_t> sj_off, sj_iph;
  std::vector<float> sj_coef;
  std::vector<float> legp;   // [npart][maxpg][ml+1]
};

void species_tables(const Plan& p, const GradPlan& g, SpeciesTables& t) {
  const int npts = p.npts, npart = p.npart, ML = p.ml, nl = p.nleg + 1, maxpg = g.maxpg;
  const int nl1 = ML + 1;
  const float scatmin = (float)p.scatmin;
  t.legt_s.assign((size_t)npart * npts * nl1, 0.f);
  t.legt_u.assign((size_t)npart * npts * nl1, 0.f);
  t.f_s.assign((size_t)npart * npts, 0.f);
  t.scatj_ok.assign((size_t)npart * npts, 0);
  t.sj_off.assign((size_t)npart * npts + 1, 0);
  t.sj_iph.clear();
  t.sj_coef.clear();
  std::vector<float> legt;
  for (int ipa = 0; ipa < npart; ipa++) {
    for (int ip = 0; ip < npts; ip++) {
      const size_t sp = (size_t)ipa * npts + ip;
      t.sj_off[sp] = (int32_t)t.sj_iph.size();
      // SCATTERJ and the SINGSCATJ pairs
      float scatterj = 0.f;
      size_t first = t.sj_iph.size();
      for (int nb = 0; nb < 8; nb++) {
        const int ib = g.interpptr[(size_t)ip * 8 + nb];
        const float xi = g.optinterpwt[(size_t)ip * 8 + nb];
        const size_t pb = (size_t)ipa * maxpg + ib;
        const float sw = xi * p.albedop[pb] * p.extinctp[pb];
        scatterj = scatterj + sw;
        if (sw <= 1e-6f) continue;
        if (p.phasewtp[pb] <= 1e-6f) continue;
        t.sj_iph.push_back(p.iphasep[pb] - 1);
        t.sj_coef.push_back(sw * p.phasewtp[pb]);
      }
      const float div = scatterj > scatmin ? scatterj : scatmin;
      for (size_t q = first; q < t.sj_coef.size(); q++) t.sj_coef[q] = t.sj_coef[q] / div;
      t.scatj_ok[sp] = scatterj > scatmin;
      float F;
      float* ls = &t.legt_s[sp * nl1];
      float* lu = &t.legt_u[sp * nl1];
      if (npart == 1) {
        plan_mixed_legen(p, ipa, ip, legt);
        F = legt[ML + 1];
        for (int l = 0; l <= ML; l++) ls[l] = legt[l] / (1 - F);
      } else {
        legt.assign(nl, 0.f);
        for (int nb = 0; nb < 8; nb++) {
          const int ib = g.interpptr[(size_t)ip * 8 + nb];
          const float xi = g.optinterpwt[(size_t)ip * 8 + nb];
          const size_t pb = (size_t)ipa * maxpg + ib;
          const float sw = xi * p.albedop[pb] * p.extinctp[pb];
          if (sw <= 1e-6f) continue;
          if (p.phasewtp[pb] <= 1e-6f) continue;
          const float* L = &p.legen[(size_t)(p.iphasep[pb] - 1) * nl];
          for (int l = 0; l < nl; l++) legt[l] = legt[l] + sw * p.phasewtp[pb] * L[l];
        }
        for (int l = 0; l < nl; l++) legt[l] = legt[l] / div;
        F = legt[ML + 1];
        for (int l = 0; l <= ML; l++) ls[l] = legt[l] / (1 - F);
      }
      t.f_s[sp] = F;
      for (int l = 0; l <= ML; l++) lu[l] = ls[l] * (1 - F) + F;
    }
  }
  t.sj_off[(size_t)npart * npts] = (int32_t)t.sj_iph.size();
  // unscaled phase function times its weight at every property point
  t.legp.assign((size_t)npart * maxpg * nl1, 0.f);
  for (int ipa = 0; ipa < npart; ipa++)
    for (int ib = 0; ib < maxpg; ib++) {
      const size_t pb = (size_t)ipa * maxpg + ib;
      const float* L = &p.legen[(size_t)(p.iphasep[pb] - 1) * nl];
      const float ft = L[ML + 1];
      for (int l = 0; l <= ML; l++) t.legp[pb * nl1 + l] = p.phasewtp[pb] * (L[l] + ft);
    }
}

}  // namespace

int state_levis_gradient(State& s, const at3d_deriv_desc& dd, int64_t nrays, const float* x,
                         const float* y, const float* z, const double* mu, const double* phi,
                         int64_t npix, const int32_t* rays_per_pixel, const double* ray_weights,
                         const double* stokes_weights, const float* measurements,
                         const double* uncertainties, const char* costfunc, int ess, int32_t njac,
                         const int32_t* jacptr, float* jacobian, double* grad, double* cost,
                         float* pixels, std::string& err) {
  const Plan& p = *s.plan;
  if (s.nst != 1) {
    err = "the Levis gradient is for NSTOKES = 1 solves";
    return AT3D_ERR_ARG;
  }
  if (p.nmicro != 1) {
    err = "the Levis gradient is for one phase table per property point (num_micro = 1)";
    return AT3D_ERR_ARG;
  }
  if (p.interp_o) {
    err = "the Levis gradient is for INTERPMETHOD 'ON' solves";
    return AT3D_ERR_ARG;
  }
  if (!s.initialized) {
    err = "the solution is not initialised: solve before computing a gradient";
    return AT3D_ERR_ARG;
  }
  {
    // the derivative set-up reads the per-point optics on the host
    const int rc = ensure_host_medium(p, err);
    if (rc) return rc;
  }
  if (!p.d.deltam) {
    err = "the gradient requires delta-M scaling (deltam = True)";
    return AT3D_ERR_ARG;
  }
  if (p.ml > AT3D_MAX_L) {
    err = "the gradient supports spherical harmonics up to L = 63";
    return AT3D_ERR_ARG;
  }
  if (!p.phasetab_err.empty()) {
    err = p.phasetab_err;
    return AT3D_ERR_NUMERIC;
  }
  const bool l2 = costfunc && strcmp(costfunc, "L2") == 0;
  const bool ll = costfunc && strcmp(costfunc, "LL") == 0;
  if (!l2 && !ll) {
    err = std::string("COMPUTE_ADJOINT_WEIGHTS: Bad cost function ") + (costfunc ? costfunc : "(null)");
    return AT3D_ERR_ARG;
  }
  if (nrays <= 0 || npix <= 0 || nrays > ((int64_t)1 << 30)) {
    err = "need at least one ray and one pixel";
    return AT3D_ERR_ARG;
  }
  {
    int64_t tot = 0;
    for (int64_t i = 0; i < npix; i++) {
      if (rays_per_pixel[i] < 0) {
        err = "rays_per_pixel must be non-negative";
        return AT3D_ERR_ARG;
      }
      tot += rays_per_pixel[i];
    }
    if (tot != nrays) {
      err = "sum of rays_per_pixel must equal the number of rays";
      return AT3D_ERR_ARG;
    }
  }
  const float zbot = p.zgrid[0];
  for (int64_t i = 0; i < nrays; i++) {
    if (z[i] < zbot) {
      err = "LEVISAPPROX_GRADIENT: Level below domain";
      return AT3D_ERR_ARG;
    }
    if (!(mu[i] >= -1.0 && mu[i] <= 1.0) || mu[i] == 0.0) {
      err = "ray mu must be in [-1, 1] and nonzero";
      return AT3D_ERR_ARG;
    }
  }
  GradPlan g;
  {
    const int rc = build_grad_plan(p, dd, g, err);
    if (rc) return rc;
  }
  // ---- Jacobian columns: the grid points whose interpolation kernel or sun
  // path reaches one of the requested property points ----
  if (njac < 0 || (njac > 0 && (!jacptr || !jacobian))) {
    err = "Jacobian: need the property point list and an output array";
    return AT3D_ERR_ARG;
  }
  std::vector<int32_t> rel_s, rel_b, p_of_rs, js_off(1, 0), js_r, js_nb, jb_off(1, 0), jb_r;
  std::vector<float> jb_w;
  int nrs = 0, nrb = 0;
  if (njac > 0) {
    for (int ji = 0; ji < njac; ji++)
      if (jacptr[ji] < 0 || jacptr[ji] >= g.maxpg) {
        err = "Jacobian: property point index out of range";
        return AT3D_ERR_ARG;
      }
    std::vector<std::vector<int32_t>> cols_of_ib(g.maxpg);
    for (int ji = 0; ji < njac; ji++) cols_of_ib[jacptr[ji]].push_back(ji);
    rel_s.assign(p.npts, -1);
    rel_b.assign(p.npts, -1);
    std::vector<std::vector<std::pair<int32_t, int32_t>>> src_of(njac);
    for (int ip = 0; ip < p.npts; ip++)
      for (int nb = 0; nb < 8; nb++)
        for (int ji : cols_of_ib[g.interpptr[(size_t)ip * 8 + nb]]) {
          if (rel_s[ip] < 0) {
            rel_s[ip] = nrs++;
            p_of_rs.push_back(ip);
          }
          src_of[ji].push_back({rel_s[ip], nb});
        }
    for (int ji = 0; ji < njac; ji++) {
      for (auto& e : src_of[ji]) {
        js_r.push_back(e.first);
        js_nb.push_back(e.second);
      }
      js_off.push_back((int32_t)js_r.size());
      const int ib = jacptr[ji];
      for (int32_t e = g.bp_off[ib]; e < g.bp_off[ib + 1]; e++) {
        const int ip = g.bp_ip[e];
        if (rel_b[ip] < 0) rel_b[ip] = nrb++;
        jb_r.push_back(rel_b[ip]);
        jb_w.push_back(g.bp_w[e]);
      }
      jb_off.push_back((int32_t)jb_r.size());
    }
    const double bytes = (double)nrays * (2.0 * nrs + nrb) * 8.0;
    if (bytes > 16.0e9) {
      err = "Jacobian: per-ray weights at the grid points near the requested property points "
            "would need more than 16 GB; request fewer points or rays per call";
      return AT3D_ERR_ARG;
    }
  }
  SpeciesTables sp;
  species_tables(p, g, sp);
  DirTables t;
  {
    const int rc = direction_tables(s, nrays, mu, phi, t, err);
    if (rc) return rc;
  }
  const int ndir = t.ndir, ldd = t.ldd, ML = p.ml, np = p.numphase, nsa = p.nscatangle;
  const int npts = p.npts, maxpg = g.maxpg, numder = g.numder, npart = p.npart, nl1 = ML + 1;
  std::vector<float> dssd((size_t)ndir * g.dnumphase, 0.f);
  for (int d = 0; d < ndir; d++) {
    const int j = t.sa_j[d];
    const float f = t.sa_f[d];
    for (int k = 0; k < g.dnumphase; k++)
      dssd[(size_t)d * g.dnumphase + k] =
          (1 - f) * g.dphasetab[(size_t)k * nsa + j - 1] + f * g.dphasetab[(size_t)k * nsa + j];
  }
  std::vector<float> dlegp((size_t)numder * maxpg * nl1, 0.f);
  for (int idr = 0; idr < numder; idr++) {
    const int ipa = g.partder[idr];
    for (int ib = 0; ib < maxpg; ib++) {
      const size_t pb = (size_t)ipa * maxpg + ib, db = (size_t)idr * maxpg + ib;
      float* o = &dlegp[db * nl1];
      if (g.doexact[idr] == 1) {
        const float* D = &g.dleg[(size_t)g.diphasep[db] * (p.nleg + 1)];
        for (int l = 0; l <= ML; l++) o[l] = p.phasewtp[pb] * D[l];
      } else {
        const float* L = &p.legen[(size_t)(p.iphasep[pb] - 1) * (p.nleg + 1)];
        const float ft = L[ML + 1];
        for (int l = 0; l <= ML; l++) o[l] = g.dphasewtp[db] * (L[l] + ft);
      }
    }
  }
  std::vector<float> topval(nrays);
  for (int64_t i = 0; i < nrays; i++) topval[i] = (mu[i] < 0.0) ? p.d.skyrad : 0.0f;
  std::vector<int32_t> partder(g.partder), doexact(g.doexact);

  // ---- device buffers (one block, grown on demand) ----
  const size_t nr = (size_t)nrays;
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    const size_t o = off;
    off = rup(off + std::max<size_t>(bytes, 4), 256);
    return o;
  };
  const size_t tab = (size_t)s.nptsp * ldd;
  const size_t o_yd = carve(t.Yd.size() * 4), o_g = carve(t.G.size() * 4), o_ss = carve(t.ssd.size() * 4);
  const size_t o_dss = carve(dssd.size() * 4);
  const size_t o_raw = carve(tab * 4), o_dsrc = carve(tab * 4), o_pre = carve(tab * 4), o_ss8 = carve(tab * 4);
  const size_t o_x = carve(nr * 4), o_y = carve(nr * 4), o_z = carve(nr * 4);
  const size_t o_mu = carve(nr * 8), o_phi = carve(nr * 8), o_dir = carve(nr * 4);
  const size_t o_top = carve(nr * 4), o_out = carve(nr * 4), o_adj = carve(nr * 8);
  const size_t o_bsrc = carve(tab * 8), o_erad = carve((size_t)npts * 8), o_bw = carve((size_t)npts * 8);
  const size_t o_grad = carve((size_t)numder * maxpg * 8);
  const size_t o_sjo = carve(sp.sj_off.size() * 4), o_sji = carve(sp.sj_iph.size() * 4);
  const size_t o_sjc = carve(sp.sj_coef.size() * 4), o_sok = carve(sp.scatj_ok.size());
  const size_t o_lts = carve(sp.legt_s.size() * 4), o_fs = carve(sp.f_s.size() * 4);
  const size_t o_ltu = carve(sp.legt_u.size() * 4), o_oiw = carve(g.optinterpwt.size() * 4);
  const size_t o_ipt = carve(g.interpptr.size() * 4), o_extp = carve(p.extinctp.size() * 4);
  const size_t o_albp = carve(p.albedop.size() * 4), o_iphp = carve(p.iphasep.size() * 4);
  const size_t o_pwp = carve(p.phasewtp.size() * 4), o_dext = carve(g.dext.size() * 4);
  const size_t o_dalb = carve(g.dalb.size() * 4), o_dpw = carve(g.dphasewtp.size() * 4);
  const size_t o_dip = carve(g.diphasep.size() * 4), o_dextm = carve(g.dextm.size() * 4);
  const size_t o_legp = carve(sp.legp.size() * 4), o_dlegp = carve(dlegp.size() * 4);
  const size_t o_dfj = carve(g.dfj.size() * 4), o_dalbm = carve(g.dalbm.size() * 4);
  const size_t o_bpo = carve(g.bp_off.size() * 4), o_bpi = carve(g.bp_ip.size() * 4);
  const size_t o_bpw = carve(g.bp_w.size() * 4), o_alb = carve(p.albedo.size() * 4);
  const size_t o_pd = carve(partder.size() * 4), o_dx = carve(doexact.size() * 4);
  // Jacobian
  const size_t o_jw = carve(njac > 0 ? nr * 8 : 0), o_rels = carve(rel_s.size() * 4);
  const size_t o_relb = carve(rel_b.size() * 4), o_prs = carve(p_of_rs.size() * 4);
  const size_t o_js = carve(nr * nrs * 8), o_jr = carve(nr * nrs * 8), o_jb = carve(nr * nrb * 8);
  const size_t o_jso = carve(js_off.size() * 4), o_jsr = carve(js_r.size() * 4);
  const size_t o_jsn = carve(js_nb.size() * 4), o_jbo = carve(jb_off.size() * 4);
  const size_t o_jbr = carve(jb_r.size() * 4), o_jbw = carve(jb_w.size() * 4);
  const size_t o_g8t = carve((size_t)nrs * ndir * numder * 8 * 4);
  const size_t o_roff = carve(njac > 0 ? ((size_t)npix + 1) * 4 : 0);
  const size_t o_jib = carve((size_t)std::max(njac, 0) * 4);
  const size_t o_jac = carve(njac > 0 ? (size_t)npix * njac * numder * 4 : 0);
  if (off > s.gbuf_bytes) {
    // the stream's earlier work may still read the old buffer
    HIPCHK(hipStreamSynchronize(s.stream));
    dev_free(s.gbuf, s.gbuf_bytes);
    s.gbuf = nullptr;
    s.gbuf_bytes = 0;
    HIPCHK(dev_alloc(&s.gbuf, off));
    s.gbuf_bytes = off;
  }
  char* B = static_cast<char*>(s.gbuf);
  auto up = [&](size_t o, const void* src, size_t bytes) -> hipError_t {
    if (bytes == 0) return hipSuccess;
    return hipMemcpyAsync(B + o, src, bytes, hipMemcpyHostToDevice, s.stream);
  };
  HIPCHK(up(o_yd, t.Yd.data(), t.Yd.size() * 4));
  HIPCHK(up(o_g, t.G.data(), t.G.size() * 4));
  HIPCHK(up(o_ss, t.ssd.data(), t.ssd.size() * 4));
  HIPCHK(up(o_dss, dssd.data(), dssd.size() * 4));
  HIPCHK(up(o_x, x, nr * 4));
  HIPCHK(up(o_y, y, nr * 4));
  HIPCHK(up(o_z, z, nr * 4));
  HIPCHK(up(o_mu, mu, nr * 8));
  HIPCHK(up(o_phi, phi, nr * 8));
  HIPCHK(up(o_dir, t.dir_of_ray.data(), nr * 4));
  HIPCHK(up(o_top, topval.data(), nr * 4));
  HIPCHK(up(o_sjo, sp.sj_off.data(), sp.sj_off.size() * 4));
  HIPCHK(up(o_sji, sp.sj_iph.data(), sp.sj_iph.size() * 4));
  HIPCHK(up(o_sjc, sp.sj_coef.data(), sp.sj_coef.size() * 4));
  HIPCHK(up(o_sok, sp.scatj_ok.data(), sp.scatj_ok.size()));
  HIPCHK(up(o_lts, sp.legt_s.data(), sp.legt_s.size() * 4));
  HIPCHK(up(o_fs, sp.f_s.data(), sp.f_s.size() * 4));
  HIPCHK(up(o_ltu, sp.legt_u.data(), sp.legt_u.size() * 4));
  HIPCHK(up(o_oiw, g.optinterpwt.data(), g.optinterpwt.size() * 4));
  HIPCHK(up(o_ipt, g.interpptr.data(), g.interpptr.size() * 4));
  HIPCHK(up(o_extp, p.extinctp.data(), p.extinctp.size() * 4));
  HIPCHK(up(o_albp, p.albedop.data(), p.albedop.size() * 4));
  HIPCHK(up(o_iphp, p.iphasep.data(), p.iphasep.size() * 4));
  HIPCHK(up(o_pwp, p.phasewtp.data(), p.phasewtp.size() * 4));
  HIPCHK(up(o_dext, g.dext.data(), g.dext.size() * 4));
  HIPCHK(up(o_dalb, g.dalb.data(), g.dalb.size() * 4));
  HIPCHK(up(o_dpw, g.dphasewtp.data(), g.dphasewtp.size() * 4));
  HIPCHK(up(o_dip, g.diphasep.data(), g.diphasep.size() * 4));
  HIPCHK(up(o_dextm, g.dextm.data(), g.dextm.size() * 4));
  HIPCHK(up(o_legp, sp.legp.data(), sp.legp.size() * 4));
  HIPCHK(up(o_dlegp, dlegp.data(), dlegp.size() * 4));
  HIPCHK(up(o_dfj, g.dfj.data(), g.dfj.size() * 4));
  HIPCHK(up(o_dalbm, g.dalbm.data(), g.dalbm.size() * 4));
  HIPCHK(up(o_bpo, g.bp_off.data(), g.bp_off.size() * 4));
  HIPCHK(up(o_bpi, g.bp_ip.data(), g.bp_ip.size() * 4));
  HIPCHK(up(o_bpw, g.bp_w.data(), g.bp_w.size() * 4));
  HIPCHK(up(o_alb, p.albedo.data(), p.albedo.size() * 4));
  HIPCHK(up(o_pd, partder.data(), partder.size() * 4));
  HIPCHK(up(o_dx, doexact.data(), doexact.size() * 4));
  std::vector<int32_t> ray_off;
  if (njac > 0) {
    HIPCHK(up(o_rels, rel_s.data(), rel_s.size() * 4));
    HIPCHK(up(o_relb, rel_b.data(), rel_b.size() * 4));
    HIPCHK(up(o_prs, p_of_rs.data(), p_of_rs.size() * 4));
    HIPCHK(up(o_jso, js_off.data(), js_off.size() * 4));
    HIPCHK(up(o_jsr, js_r.data(), js_r.size() * 4));
    HIPCHK(up(o_jsn, js_nb.data(), js_nb.size() * 4));
    HIPCHK(up(o_jbo, jb_off.data(), jb_off.size() * 4));
    HIPCHK(up(o_jbr, jb_r.data(), jb_r.size() * 4));
    HIPCHK(up(o_jbw, jb_w.data(), jb_w.size() * 4));
    ray_off.assign(npix + 1, 0);
    for (int64_t i = 0; i < npix; i++) ray_off[i + 1] = ray_off[i] + rays_per_pixel[i];
    HIPCHK(up(o_roff, ray_off.data(), ray_off.size() * 4));
    HIPCHK(up(o_jib, jacptr, (size_t)njac * 4));
    HIPCHK(hipMemsetAsync(B + o_js, 0, nr * nrs * 8, s.stream));
    HIPCHK(hipMemsetAsync(B + o_jr, 0, nr * nrs * 8, s.stream));
    HIPCHK(hipMemsetAsync(B + o_jb, 0, nr * nrb * 8, s.stream));
  }
  HIPCHK(hipMemsetAsync(s.errflag, 0, sizeof(int), s.stream));

  // ---- phase 1: forward radiances (INTEGRATE_1RAY) ----
  float* raw = reinterpret_cast<float*>(B + o_raw);
  float* dsrc = reinterpret_cast<float*>(B + o_dsrc);
  float* srcpre = reinterpret_cast<float*>(B + o_pre);
  float* ss8 = reinterpret_cast<float*>(B + o_ss8);
  launch_sh_dot(s.stream, s.S, reinterpret_cast<float*>(B + o_yd), raw, s.nptsp, ldd, s.ld, ldd, npts, ndir);
  launch_dir_source_grad(s.stream, raw, ldd, ndir, s.ext, s.ns, s.lofj, s.tms,
                         reinterpret_cast<float*>(B + o_g), s.ss_off, s.ss_iph, s.ss_coef,
                         reinterpret_cast<float*>(B + o_ss), np, ML, npts, 0, dsrc, srcpre, ss8);
  const float albpi = p.d.gndalbedo / acosf(-1.0f);
  launch_render_bottom(s.stream, s.fluxes, s.dirflux, s.bcbot, (int)p.bcbot.size(), albpi, s.bcrad_render);
  RayArgs a;
  a.g = GridTopo{s.gridpos, s.gridptr, s.neighptr, s.cellflags, p.npts, p.ncells, s.treeptr};
  a.xgrid = s.xgrid;
  a.ygrid = s.ygrid;
  a.zgrid = s.zgrid;
  a.nx = p.nx;
  a.ny = p.ny;
  a.nz = p.nz;
  a.nx1 = p.nx1;
  a.ny1 = p.ny1;
  a.nyc = p.nyc;
  a.open_x = p.bcflag & 1;
  a.open_y = (p.bcflag >> 1) & 1;
  a.ext = s.ext;
  a.dsrc = dsrc;
  a.ldd = ldd;
  a.bot_index = s.bot_index;
  a.bcrad_bot = s.bcrad_render;
  a.camx = reinterpret_cast<float*>(B + o_x);
  a.camy = reinterpret_cast<float*>(B + o_y);
  a.camz = reinterpret_cast<float*>(B + o_z);
  a.cammu = reinterpret_cast<double*>(B + o_mu);
  a.camphi = reinterpret_cast<double*>(B + o_phi);
  a.dir_of_ray = reinterpret_cast<int32_t*>(B + o_dir);
  a.topval = reinterpret_cast<float*>(B + o_top);
  a.nrays = (int)nrays;
  a.tautol = p.d.tautol;
  a.transcut = p.d.transcut;
  a.nosurface = 0;
  a.out = reinterpret_cast<float*>(B + o_out);
  a.errflag = s.errflag;
  (void)hipEventRecord(s.ev[6], s.stream);
  launch_render(s.stream, a);
  HIPCHK(hipGetLastError());
  std::vector<float> vis(nr);
  int ef = 0;
  HIPCHK(hipMemcpyAsync(vis.data(), B + o_out, nr * 4, hipMemcpyDeviceToHost, s.stream));
  HIPCHK(hipMemcpyAsync(&ef, s.errflag, sizeof(int), hipMemcpyDeviceToHost, s.stream));
  HIPCHK(hipStreamSynchronize(s.stream));
  if (ef) {
    err = (ef & 1)   ? "INTEGRATE_1RAY: SO<0"
          : (ef & 2) ? "FIND_BOUNDARY_RADIANCE: Not at boundary"
                     : "LOCATE_GRID_CELL: ray start outside the grid";
    return AT3D_ERR_NUMERIC;
  }

  // ---- phase 2: pixels, cost and adjoint weights (COMPUTE_ADJOINT_WEIGHTS) ----
  std::vector<double> adjw_pix(npix, 0.0), adjw_ray(nr, 0.0), jacw_ray(njac > 0 ? nr : 0, 0.0);
  double c = 0.0;
  {
    int64_t iray = 0;
    for (int64_t ip = 0; ip < npix; ip++) {
      float so = 0.f;
      for (int k = 0; k < rays_per_pixel[ip]; k++, iray++)
        so = (float)((double)so + (double)vis[iray] * ray_weights[iray] * stokes_weights[ip]);
      pixels[ip] = so;
      const double r = (double)so, m = (double)measurements[ip], u = uncertainties[ip];
      if (l2) {
        const double e = r - m;
        c = c + 0.5 * u * e * e;
        adjw_pix[ip] = u * e;
      } else {
        const double e = log(r) - log(m);
        c = c + 0.5 * (e * e * u);
        adjw_pix[ip] = e * u / r;
      }
    }
    iray = 0;
    for (int64_t ip = 0; ip < npix; ip++)
      for (int k = 0; k < rays_per_pixel[ip]; k++, iray++) {
        adjw_ray[iray] = adjw_pix[ip] * ray_weights[iray] * stokes_weights[ip];
        if (njac > 0) jacw_ray[iray] = ray_weights[iray] * stokes_weights[ip];
      }
  }
  *cost = c;

  // ---- phase 3: adjoint traversal and contraction ----
  HIPCHK(up(o_adj, adjw_ray.data(), nr * 8));
  if (njac > 0) HIPCHK(up(o_jw, jacw_ray.data(), nr * 8));
  HIPCHK(hipMemsetAsync(B + o_bsrc, 0, tab * 8, s.stream));
  HIPCHK(hipMemsetAsync(B + o_erad, 0, (size_t)npts * 8, s.stream));
  HIPCHK(hipMemsetAsync(B + o_bw, 0, (size_t)npts * 8, s.stream));
  HIPCHK(hipMemsetAsync(B + o_grad, 0, (size_t)numder * maxpg * 8, s.stream));
  AdjArgs A;
  A.r = a;
  A.ss8 = ss8;
  A.adj_weight = reinterpret_cast<double*>(B + o_adj);
  A.dirflux = s.dirflux;
  A.opi_gndalbedo = (1.0f / acosf(-1.0f)) * p.d.gndalbedo;
  A.exact_single_scatter = ess;
  A.max_intervals = std::max(50 * std::max(p.nx, std::max(p.ny, p.nz)), s.max_subgrid);
  A.bsrc = reinterpret_cast<double*>(B + o_bsrc);
  A.erad = reinterpret_cast<double*>(B + o_erad);
  A.beamw = reinterpret_cast<double*>(B + o_bw);
  A.jac_weight = njac > 0 ? reinterpret_cast<double*>(B + o_jw) : nullptr;
  A.rel_s = reinterpret_cast<int32_t*>(B + o_rels);
  A.rel_b = reinterpret_cast<int32_t*>(B + o_relb);
  A.nrel_s = nrs;
  A.nrel_b = nrb;
  A.jsrc = reinterpret_cast<double*>(B + o_js);
  A.jrad = reinterpret_cast<double*>(B + o_jr);
  A.jbeam = reinterpret_cast<double*>(B + o_jb);
  (void)hipEventRecord(s.ev_ray[0], s.stream);
  launch_adjoint_rays(s.stream, A);
  (void)hipEventRecord(s.ev_ray[1], s.stream);
  GradPointArgs G;
  G.npts = npts;
  G.ndir = ndir;
  G.ldd = ldd;
  G.ld = s.ld;
  G.ml = ML;
  G.numder = numder;
  G.npart = npart;
  G.maxpg = maxpg;
  G.numphase = np;
  G.dnumphase = g.dnumphase;
  G.secmu0 = (float)(1.0 / fabs((double)p.d.solarmu));
  G.bsrc = A.bsrc;
  G.erad = A.erad;
  G.beamw = A.beamw;
  G.grad = reinterpret_cast<double*>(B + o_grad);
  G.R = s.R;
  G.Yd = reinterpret_cast<float*>(B + o_yd);
  G.nr = s.nr;
  G.lofj = s.lofj;
  G.dirflux = s.dirflux;
  G.ssd = reinterpret_cast<float*>(B + o_ss);
  G.dssd = reinterpret_cast<float*>(B + o_dss);
  G.partder = reinterpret_cast<int32_t*>(B + o_pd);
  G.doexact = reinterpret_cast<int32_t*>(B + o_dx);
  G.sj_off = reinterpret_cast<int32_t*>(B + o_sjo);
  G.sj_iph = reinterpret_cast<int32_t*>(B + o_sji);
  G.sj_coef = reinterpret_cast<float*>(B + o_sjc);
  G.albedo = reinterpret_cast<float*>(B + o_alb);
  G.srcpre = srcpre;
  G.scatj_ok = reinterpret_cast<int8_t*>(B + o_sok);
  G.legt_s = reinterpret_cast<float*>(B + o_lts);
  G.f_s = reinterpret_cast<float*>(B + o_fs);
  G.legt_u = reinterpret_cast<float*>(B + o_ltu);
  G.optinterpwt = reinterpret_cast<float*>(B + o_oiw);
  G.interpptr = reinterpret_cast<int32_t*>(B + o_ipt);
  G.extinctp = reinterpret_cast<float*>(B + o_extp);
  G.albedop = reinterpret_cast<float*>(B + o_albp);
  G.iphasep = reinterpret_cast<int32_t*>(B + o_iphp);
  G.phasewtp = reinterpret_cast<float*>(B + o_pwp);
  G.dext = reinterpret_cast<float*>(B + o_dext);
  G.dalb = reinterpret_cast<float*>(B + o_dalb);
  G.dphasewtp = reinterpret_cast<float*>(B + o_dpw);
  G.diphasep = reinterpret_cast<int32_t*>(B + o_dip);
  G.dextm = reinterpret_cast<float*>(B + o_dextm);
  G.legp = reinterpret_cast<float*>(B + o_legp);
  G.dlegp = reinterpret_cast<float*>(B + o_dlegp);
  G.dfj = reinterpret_cast<float*>(B + o_dfj);
  G.dalbm = reinterpret_cast<float*>(B + o_dalbm);
  G.bp_off = reinterpret_cast<int32_t*>(B + o_bpo);
  G.bp_ip = reinterpret_cast<int32_t*>(B + o_bpi);
  G.bp_w = reinterpret_cast<float*>(B + o_bpw);
  launch_grad_points(s.stream, G);
  if (ess) launch_grad_beam(s.stream, G);
  if (njac > 0) {
    JacArgs J;
    J.npix = (int)npix;
    J.njac = njac;
    J.numder = numder;
    J.ndir = ndir;
    J.maxpg = maxpg;
    J.nrel_s = nrs;
    J.nrel_b = nrb;
    J.ess = ess;
    J.ray_off = reinterpret_cast<int32_t*>(B + o_roff);
    J.dir_of_ray = a.dir_of_ray;
    J.jac_ib = reinterpret_cast<int32_t*>(B + o_jib);
    J.js_off = reinterpret_cast<int32_t*>(B + o_jso);
    J.js_r = reinterpret_cast<int32_t*>(B + o_jsr);
    J.js_nb = reinterpret_cast<int32_t*>(B + o_jsn);
    J.jb_off = reinterpret_cast<int32_t*>(B + o_jbo);
    J.jb_r = reinterpret_cast<int32_t*>(B + o_jbr);
    J.jb_w = reinterpret_cast<float*>(B + o_jbw);
    J.p_of_rs = reinterpret_cast<int32_t*>(B + o_prs);
    J.g8t = reinterpret_cast<float*>(B + o_g8t);
    J.dextm = G.dextm;
    J.optinterpwt = G.optinterpwt;
    J.jsrc = A.jsrc;
    J.jrad = A.jrad;
    J.jbeam = A.jbeam;
    J.jacobian = reinterpret_cast<float*>(B + o_jac);
    launch_jacobian(s.stream, G, J, reinterpret_cast<float*>(B + o_g8t));
  }
  (void)hipEventRecord(s.ev[7], s.stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(grad, B + o_grad, (size_t)numder * maxpg * 8, hipMemcpyDeviceToHost, s.stream));
  if (njac > 0)
    HIPCHK(hipMemcpyAsync(jacobian, B + o_jac, (size_t)npix * njac * numder * 4, hipMemcpyDeviceToHost,
                          s.stream));
  HIPCHK(hipMemcpyAsync(&ef, s.errflag, sizeof(int), hipMemcpyDeviceToHost, s.stream));
  HIPCHK(hipStreamSynchronize(s.stream));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, s.ev[6], s.ev[7]);
  s.grad_ms = ms;
  (void)hipEventElapsedTime(&ms, s.ev_ray[0], s.ev_ray[1]);
  s.adjoint_ms = ms;
  if (ef) {
    err = (ef & 32)  ? "ADJOINT_INTEGRATE_1RAY: ray crossed too many cells without leaving the grid"
          : (ef & 16) ? "ADJOINT_INTEGRATE_1RAY: The maximum number of subgrid intervals for calculation "
                        "of the radiance along the ray path has been exceeded."
          : (ef & 1) ? "ADJOINT_INTEGRATE_1RAY: SO<0"
          : (ef & 2) ? "FIND_BOUNDARY_RADIANCE: Not at boundary"
                     : "LOCATE_GRID_CELL: ray start outside the grid";
    return AT3D_ERR_NUMERIC;
  }
  return AT3D_OK;
}

}  // namespace at3d
