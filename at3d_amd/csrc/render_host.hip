// Host orchestration of rendering (RTE.integrate_to_sensor -> RENDER,
// shdomsub4.f:93-286): per distinct ray direction the spherical harmonics and
// the single-scatter phase values, then the direction source for every grid
// point on the GPU and one ray per lane.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

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

static size_t rup(size_t v, size_t m) { return (v + m - 1) / m * m; }

// Distinct ray directions and their tables: spherical harmonics Y_J(dir)
// [ldd][ld], the truncated single-scatter factors G[dir][l] = sum over J of
// degree l of Y_J(dir) Y_J(sun), and the tabulated phase functions at the
// solar scattering angle [dir][numphase] (linear in angle, as
// INTEGRATE_1RAY shdomsub2.f:2393-2433).
int direction_tables(const State& s, int64_t nrays, const double* mu, const double* phi, DirTables& t,
                     std::string& err) {
  const Plan& p = *s.plan;
  std::map<std::pair<double, double>, int> dmap;
  t.dir_of_ray.assign(nrays, 0);
  t.dirs.clear();
  for (int64_t i = 0; i < nrays; i++) {
    auto key = std::make_pair(mu[i], phi[i]);
    auto it = dmap.find(key);
    if (it == dmap.end()) {
      it = dmap.emplace(key, (int)t.dirs.size()).first;
      t.dirs.push_back(key);
    }
    t.dir_of_ray[i] = it->second;
  }
  const int ndir = (int)t.dirs.size();
  const int ldd = (int)rup(ndir, 64);
  t.ndir = ndir;
  t.ldd = ldd;
  if ((int64_t)s.nptsp * ldd >= ((int64_t)1 << 31)) {
    err = "too many distinct ray directions for the direction-source table";
    return AT3D_ERR_ARG;
  }
  const int ML = p.ml, nlm = p.nlm, np = p.numphase, nsa = p.nscatangle;
  const bool tms_on = p.d.deltam != 0;
  const bool pol = p.nst == 3;
  const int ld = s.ld;
  t.Yd.assign((size_t)ldd * ld, 0.f);
  t.G.assign((size_t)(pol ? 3 : 1) * ndir * (ML + 1), 0.f);
  t.ssd.assign((size_t)ndir * np, 0.f);
  if (pol) {
    t.Yqu.assign((size_t)4 * ldd * ld, 0.f);
    t.ssq.assign((size_t)ndir * np, 0.f);
    t.ssu.assign((size_t)ndir * np, 0.f);
  }
  t.sa_j.assign(ndir, 1);
  t.sa_f.assign(ndir, 0.f);
  std::vector<float> ylm(nlm), y6(pol ? (size_t)nlm * 6 : 0);
  const double PI = acos(-1.0);
  const float solarmu = p.d.solarmu, solaraz = p.d.solaraz;
  for (int d = 0; d < ndir; d++) {
    const double mu2 = t.dirs[d].first, phi2 = t.dirs[d].second;
    if (pol) {
      // YLMALL with NSTLEG = 6 (INTEGRATE_1RAY, shdomsub2.f:2411): element 1
      // for I; Q from elements 2 (S^Q) and 5 (S^U), U from 6 and 3
      ylm_direction_pol((float)mu2, (float)phi2, ML, p.mm, y6.data());
      for (int j = 0; j < nlm; j++) {
        ylm[j] = y6[(size_t)j * 6];
        float* q = &t.Yqu[(size_t)d * 2 * ld];
        float* u = &t.Yqu[(size_t)(ldd + d) * 2 * ld];
        q[j] = y6[(size_t)j * 6 + 1];
        q[ld + j] = y6[(size_t)j * 6 + 4];
        u[j] = y6[(size_t)j * 6 + 5];
        u[ld + j] = y6[(size_t)j * 6 + 2];
      }
      // the truncated single scatter of Q and U (b1 with elements 2 and 6
      // of the direction and element 1 of the sun)
      std::vector<double> gq(ML + 1, 0.0), gu(ML + 1, 0.0);
      for (int j = 0; j < nlm; j++) {
        gq[p.lofj[j]] += (double)y6[(size_t)j * 6 + 1] * (double)p.ylmsun[j];
        gu[p.lofj[j]] += (double)y6[(size_t)j * 6 + 5] * (double)p.ylmsun[j];
      }
      for (int l = 0; l <= ML; l++) {
        t.G[(size_t)(ndir + d) * (ML + 1) + l] = (float)gq[l];
        t.G[(size_t)(2 * ndir + d) * (ML + 1) + l] = (float)gu[l];
      }
    } else {
      ylm_direction((float)mu2, (float)phi2, ML, p.mm, ylm.data());
    }
    for (int j = 0; j < nlm; j++) t.Yd[(size_t)d * ld + j] = ylm[j];
    std::vector<double> g(ML + 1, 0.0);
    for (int j = 0; j < nlm; j++) g[p.lofj[j]] += (double)ylm[j] * (double)p.ylmsun[j];
    for (int l = 0; l <= ML; l++) t.G[(size_t)d * (ML + 1) + l] = (float)g[l];
    if (tms_on) {
      double cosscat = (double)solarmu * mu2 +
                       sqrt((double)(1.0f - solarmu * solarmu) * (1.0 - mu2 * mu2)) *
                           cos((double)solaraz - phi2);
      cosscat = std::max(std::min(1.0, cosscat), -1.0);
      float f = (float)((nsa - 1) * (acos(cosscat) / PI) + 1);
      const int j = std::min(nsa - 1, (int)f);
      f = f - (float)j;
      t.sa_j[d] = j;
      t.sa_f[d] = f;
      for (int iph = 0; iph < np; iph++)
        t.ssd[(size_t)d * np + iph] =
            (1 - f) * p.phasetab[(size_t)iph * nsa + j - 1] + f * p.phasetab[(size_t)iph * nsa + j];
      if (pol) {
        // ROTATE_POL_PLANE (shdomsub2.f:3277-3316): the I-Q element from the
        // scattering plane into the outgoing meridian plane
        const float mu = (float)mu2, delphi = solaraz - (float)phi2;
        const double sin_scat = sqrt(std::max(0.0, 1.0 - cosscat * cosscat));
        const double sin_t1 = sqrt(1.0 - (double)(solarmu * solarmu));
        const double sin_t2 = sqrt(1.0 - (double)(mu * mu));
        const double sinphi = sin((double)delphi), cosphi = cos((double)delphi);
        double sin2 = 0.0, cos2 = -1.0;
        if (sin_scat != 0.0) {
          sin2 = sin_t1 * sinphi / sin_scat;
          cos2 = (sin_t2 * (double)solarmu - sin_t1 * (double)mu * cosphi) / sin_scat;
        }
        const double sin22 = 2.0 * sin2 * cos2, cos22 = 1.0 - 2.0 * sin2 * sin2;
        for (int iph = 0; iph < np; iph++) {
          const float b1 =
              (1 - f) * p.phasetab2[(size_t)iph * nsa + j - 1] + f * p.phasetab2[(size_t)iph * nsa + j];
          t.ssq[(size_t)d * np + iph] = (float)((double)b1 * cos22);
          t.ssu[(size_t)d * np + iph] = (float)((double)b1 * sin22);
        }
      }
    }
  }
  return AT3D_OK;
}

// The render body.  Its device tables (direction harmonics and factors, the
// direction source of every grid point, the rays and their radiances) are
// carved from one block: the state's render buffer, grown on demand, or, when
// `fixed` is given, that block of `fixed_cap` bytes and nothing else.
static int render_impl(State& s, int64_t nrays, const float* x, const float* y, const float* z,
                       const double* mu, const double* phi, int single_scatter, int nosurface,
                       float* out, void* fixed, size_t fixed_cap, std::string& err) {
  const Plan& p = *s.plan;
  if (nrays <= 0) return AT3D_OK;
  if (nrays > (int64_t)1 << 30) {
    err = "too many rays in one call";
    return AT3D_ERR_ARG;
  }
  if (!p.phasetab_err.empty()) {
    err = p.phasetab_err;
    return AT3D_ERR_NUMERIC;
  }
  const float ztop = p.zgrid[p.nz - 1], zbot = p.zgrid[0];
  for (int64_t i = 0; i < nrays; i++) {
    if (z[i] < zbot) {
      err = "RENDER: Level below domain";
      return AT3D_ERR_ARG;
    }
    if (!(mu[i] >= -1.0 && mu[i] <= 1.0) || mu[i] == 0.0) {
      err = "ray mu must be in [-1, 1] and nonzero";
      return AT3D_ERR_ARG;
    }
  }
  DirTables t;
  {
    const int rc = direction_tables(s, nrays, mu, phi, t, err);
    if (rc) return rc;
  }
  const int ndir = t.ndir, ldd = t.ldd;
  const int ML = p.ml, np = p.numphase;
  const bool tms_on = p.d.deltam != 0;
  // polarized: the direction source rows hold I, Q and U blocks of ldd
  // directions; the rays' Stokes vectors come back as [nrays][3]
  const int nst = p.nst, ldr = nst * ldd;
  std::vector<int32_t>& dir_of_ray = t.dir_of_ray;
  std::vector<float>& Yd = t.Yd;
  std::vector<float>& G = t.G;
  std::vector<float>& ssd = t.ssd;
  std::vector<float> topval(nrays);
  for (int64_t i = 0; i < nrays; i++) topval[i] = (mu[i] < 0.0) ? p.d.skyrad : 0.0f;
  // device buffers in one block
  const size_t nr = (size_t)nrays;
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    const size_t o = off;
    off = rup(off + bytes, 256);
    return o;
  };
  const size_t o_yd = carve(Yd.size() * 4), o_g = carve(G.size() * 4), o_ss = carve(ssd.size() * 4 + 4);
  const size_t o_yqu = carve(t.Yqu.size() * 4), o_sq = carve(t.ssq.size() * 4 + 4), o_su = carve(t.ssu.size() * 4 + 4);
  const size_t o_raw = carve((size_t)s.nptsp * ldr * 4);
  const size_t o_x = carve(nr * 4), o_y = carve(nr * 4), o_z = carve(nr * 4);
  const size_t o_mu = carve(nr * 8), o_phi = carve(nr * 8), o_dir = carve(nr * 4);
  const size_t o_top = carve(nr * 4), o_out = carve(nr * 4 * nst);
  char* B = nullptr;
  if (fixed) {
    if (off > fixed_cap) {
      char buf[200];
      snprintf(buf, sizeof buf,
               "render needs %zu bytes of device scratch (%d grid points x %d directions, %lld rays); "
               "the allocation-free entry point holds %zu",
               off, p.npts, ndir, (long long)nrays, fixed_cap);
      err = buf;
      return AT3D_ERR_ARG;
    }
    B = static_cast<char*>(fixed);
  } else {
    if (off > s.rbuf_bytes) {
      // the stream's earlier work may still read the old buffer
      HIPCHK(hipStreamSynchronize(s.stream));
      dev_free(s.rbuf, s.rbuf_bytes);
      s.rbuf = nullptr;
      s.rbuf_bytes = 0;
      HIPCHK(dev_alloc(&s.rbuf, off));
      s.rbuf_bytes = off;
    }
    B = static_cast<char*>(s.rbuf);
  }
  HIPCHK(hipMemcpyAsync(B + o_yd, Yd.data(), Yd.size() * 4, hipMemcpyHostToDevice, s.stream));
  HIPCHK(hipMemcpyAsync(B + o_g, G.data(), G.size() * 4, hipMemcpyHostToDevice, s.stream));
  if (!ssd.empty())
    HIPCHK(hipMemcpyAsync(B + o_ss, ssd.data(), ssd.size() * 4, hipMemcpyHostToDevice, s.stream));
  if (nst == 3) {
    HIPCHK(hipMemcpyAsync(B + o_yqu, t.Yqu.data(), t.Yqu.size() * 4, hipMemcpyHostToDevice, s.stream));
    HIPCHK(hipMemcpyAsync(B + o_sq, t.ssq.data(), t.ssq.size() * 4, hipMemcpyHostToDevice, s.stream));
    HIPCHK(hipMemcpyAsync(B + o_su, t.ssu.data(), t.ssu.size() * 4, hipMemcpyHostToDevice, s.stream));
  }
  HIPCHK(hipMemcpyAsync(B + o_x, x, nr * 4, hipMemcpyHostToDevice, s.stream));
  HIPCHK(hipMemcpyAsync(B + o_y, y, nr * 4, hipMemcpyHostToDevice, s.stream));
  HIPCHK(hipMemcpyAsync(B + o_z, z, nr * 4, hipMemcpyHostToDevice, s.stream));
  HIPCHK(hipMemcpyAsync(B + o_mu, mu, nr * 8, hipMemcpyHostToDevice, s.stream));
  HIPCHK(hipMemcpyAsync(B + o_phi, phi, nr * 8, hipMemcpyHostToDevice, s.stream));
  HIPCHK(hipMemcpyAsync(B + o_dir, dir_of_ray.data(), nr * 4, hipMemcpyHostToDevice, s.stream));
  HIPCHK(hipMemcpyAsync(B + o_top, topval.data(), nr * 4, hipMemcpyHostToDevice, s.stream));
  HIPCHK(hipMemsetAsync(s.errflag, 0, sizeof(int), s.stream));
  (void)hipEventRecord(s.ev[6], s.stream);
  float* raw = reinterpret_cast<float*>(B + o_raw);
  launch_sh_dot(s.stream, s.S, reinterpret_cast<float*>(B + o_yd), raw, s.nptsp, ldd, s.ld, ldr,
                p.npts, ndir);
  if (nst == 3) {
    launch_sh_dot_qu(s.stream, s.S + s.psh, s.psh, reinterpret_cast<float*>(B + o_yqu), raw + ldd, s.nptsp,
                     2 * ldd, s.ld, ldr, p.npts);
    launch_dir_source_finish_pol(s.stream, raw, ldr, ldd, ndir, s.ext, s.ns, s.lofj, s.tms, s.tms5,
                                 reinterpret_cast<float*>(B + o_g), s.ss_off, s.ss_iph, s.ss_coef,
                                 reinterpret_cast<float*>(B + o_ss), reinterpret_cast<float*>(B + o_sq),
                                 reinterpret_cast<float*>(B + o_su), np, ML, p.npts, single_scatter,
                                 tms_on ? 1 : 0, raw);
  } else {
    launch_dir_source_finish(s.stream, raw, ldd, ndir, s.ext, s.ns, s.lofj, s.tms,
                             reinterpret_cast<float*>(B + o_g), s.ss_off, s.ss_iph, s.ss_coef,
                             reinterpret_cast<float*>(B + o_ss), np, ML, p.npts, single_scatter,
                             tms_on ? 1 : 0, raw);
  }
  const float albpi = p.d.gndalbedo / acosf(-1.0f);
  launch_render_bottom(s.stream, s.fluxes, s.dirflux, s.bcbot, (int)p.bcbot.size(), albpi,
                       s.bcrad_render);
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
  a.dsrc = raw;
  a.ldd = ldr;
  a.nst = nst;
  a.cst = ldd;
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
  a.nosurface = nosurface;
  a.correct_interp = s.render_correct_interp;
  a.out = reinterpret_cast<float*>(B + o_out);
  a.errflag = s.errflag;
  (void)hipEventRecord(s.ev_ray[0], s.stream);
  launch_render(s.stream, a);
  (void)hipEventRecord(s.ev_ray[1], s.stream);
  (void)hipEventRecord(s.ev[7], s.stream);
  HIPCHK(hipGetLastError());
  int ef = 0;
  HIPCHK(hipMemcpyAsync(out, B + o_out, nr * 4 * nst, hipMemcpyDeviceToHost, s.stream));
  HIPCHK(hipMemcpyAsync(&ef, s.errflag, sizeof(int), hipMemcpyDeviceToHost, s.stream));
  HIPCHK(hipStreamSynchronize(s.stream));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, s.ev[6], s.ev[7]);
  s.render_ms = ms;
  (void)hipEventElapsedTime(&ms, s.ev_ray[0], s.ev_ray[1]);
  s.render_ray_ms = ms;
  if (ef) {
    err = (ef & 1)   ? "INTEGRATE_1RAY: SO<0"
          : (ef & 2) ? "FIND_BOUNDARY_RADIANCE: Not at boundary"
                     : "LOCATE_GRID_CELL: ray start outside the grid";
    return AT3D_ERR_NUMERIC;
  }
  return AT3D_OK;
}

int state_render(State& s, int64_t nrays, const float* x, const float* y, const float* z,
                 const double* mu, const double* phi, int single_scatter, int nosurface,
                 float* out, std::string& err) {
  return render_impl(s, nrays, x, y, z, mu, phi, single_scatter, nosurface, out, nullptr, 0, err);
}

// Device scratch of the allocation-free render (at3d_render_noalloc): a
// file-scope array every GPU holds a copy of, at its own address, from the
// library's load until the process ends.  Host code never passes `scratch`
// itself to a kernel or copy (its host name is only a stand-in): the device
// address of each GPU's copy is read once, with that GPU current, and kept.
static __device__ __align__(256) unsigned char scratch[RENDER_SCRATCH_BYTES];

static std::mutex scratch_mutex;
static std::vector<void*> scratch_ptr;   // per device ordinal

static int scratch_address(int device, void** ptr, std::string& err) {
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if ((int)scratch_ptr.size() <= device) scratch_ptr.resize(device + 1, nullptr);
  if (!scratch_ptr[device]) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != device) HIPCHK(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) HIPCHK(hipSetDevice(cur));
    if (e != hipSuccess || !p) {
      err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return AT3D_ERR_DEVICE;
    }
    scratch_ptr[device] = p;
  }
  *ptr = scratch_ptr[device];
  return AT3D_OK;
}

int state_render_noalloc(State& s, int64_t nrays, const float* x, const float* y, const float* z,
                         const double* mu, const double* phi, int single_scatter, int nosurface,
                         float* out, std::string& err) {
  void* buf = nullptr;
  const int rc = scratch_address(s.device, &buf, err);
  if (rc) return rc;
  return render_impl(s, nrays, x, y, z, mu, phi, single_scatter, nosurface, out, buf,
                     RENDER_SCRATCH_BYTES, err);
}

}  // namespace at3d
