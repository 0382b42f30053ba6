// Solution transfer between host arrays and a device state, and the
// spherical-harmonic diagnostics of a solution.
//
// Host format: the reference's ragged spherical-harmonic arrays
// (RTE.save_solution / RTE.load_solution, at3d/solver.py:1519-1686):
//   SHPTR[npts+1], SOURCE[SHPTR[npts]]   (point i holds SHPTR[i+1]-SHPTR[i] terms)
//   RSHPTR[npts+2], RADIANCE[RSHPTR[npts]]  (RSHPTR[npts+1] = RSHPTR[npts])
//   FLUXES[2][npts]                      (down, up)
// Device format: dense rows [nptsp][ld] with per-point term counts, zero
// beyond each count (the sweep, transforms and source pass rely on it).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>

#include <algorithm>
#include <string>
#include <vector>

#include "devmem.h"
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

// Ragged rows -> dense rows: one thread per (point, term slot); terms past a
// point's count (and the padding rows) are written as zero.
__global__ __launch_bounds__(256) void k_expand_sh(const float* __restrict__ src,
                                                   const int32_t* __restrict__ ptr, int npts,
                                                   int nptsp, int ld, float* __restrict__ dst,
                                                   int* __restrict__ count) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)nptsp * ld) return;
  const int i = (int)(idx / ld), j = (int)(idx - (size_t)i * ld);
  float v = 0.f;
  if (i < npts) {
    const int b = ptr[i], n = ptr[i + 1] - b;
    if (j < n) v = src[b + j];
    if (j == 0 && count) count[i] = n;
  }
  dst[idx] = v;
}

// COMPUTE_SH (shdomsub2.f:1770-1845) and COMPUTE_NETFLUXDIV (:1725-1765) for a
// solar source, one thread per point.  R is the dense radiance, nr its term
// counts; sh[k][i] for k < nshout (mean radiance, Fx, Fy, Fz, and with
// nshout 5 the rms of the terms J >= 5 including the direct beam's), and
// the net flux divergence -(sqrt(4 pi) a R_1 + a F_dir / |mu0|) with the
// absorption coefficient a = ext (1 - albedo).
__global__ __launch_bounds__(256) void k_sh_terms(const float* __restrict__ R, int ld,
                                                  const int* __restrict__ nr,
                                                  const float* __restrict__ dirflux,
                                                  const float* __restrict__ ysun,
                                                  const float* __restrict__ ext,
                                                  const float* __restrict__ alb, int npts, int mm,
                                                  float solarmu, float solaraz, int nshout,
                                                  float* __restrict__ sh, float* __restrict__ nfd) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= npts) return;
  const float PI = acosf(-1.0f);
  const float A = sqrtf(2.0f * PI / 3.0f), C = sqrtf(4.0f * PI / 3.0f), C0 = 1.0f / sqrtf(4.0f * PI);
  const float secsol = 1.0f / fabsf(solarmu);
  const float sinsol = sqrtf(1.0f - solarmu * solarmu);
  const float solx = sinsol * cosf(solaraz) * secsol, soly = sinsol * sinf(solaraz) * secsol;
  const float solm = secsol / (4.0f * PI);
  const float* r = R + (size_t)i * ld;
  const float df = dirflux[i];
  float imean = C0 * r[0];
  float fz = C * r[2];
  float fx = 0.f, fy = 0.f;
  if (mm > 0) {
    fx = -A * (r[3] + r[1]);
    fy = -A * (r[1] - r[3]);
  }
  imean = imean + df * solm;
  fx = fx + df * solx;
  fy = fy + df * soly;
  fz = fz - df;
  sh[i] = imean;
  sh[(size_t)npts + i] = fx;
  sh[(size_t)2 * npts + i] = fy;
  sh[(size_t)3 * npts + i] = fz;
  if (nshout == 5) {
    const float f0 = secsol * df;
    float horms = 0.f;
    for (int j = 4; j < nr[i]; j++) {
      const float t = f0 * ysun[j] + r[j];
      horms = horms + t * t;
    }
    sh[(size_t)4 * npts + i] = sqrtf(horms);
  }
  const float absorb = ext[i] * (1.0f - alb[i]);
  const float c4 = sqrtf(4.0f * PI);
  float hr = c4 * absorb * r[0];
  hr = hr + absorb * df * secsol;
  nfd[i] = -hr;
}

static int check_ptr(const int32_t* ptr, int npts, int nmax, int nmin, const char* what,
                     std::string& err) {
  if (ptr[0] != 0) {
    err = std::string(what) + "[0] must be 0";
    return AT3D_ERR_ARG;
  }
  for (int i = 0; i < npts; i++) {
    const int n = ptr[i + 1] - ptr[i];
    if (n < nmin || n > nmax) {
      char buf[160];
      snprintf(buf, sizeof buf, "%s: point %d holds %d terms (must be %d..%d)", what, i, n, nmin, nmax);
      err = buf;
      return AT3D_ERR_ARG;
    }
  }
  return AT3D_OK;
}

// A loaded solution (INIT_SOLUTION with INRADFLAG false, shdomsub1.f:356, as
// RTE._init_solution does after load_solution, at3d/solver.py:2654-2664): the
// source and radiance are taken as given (no two-stream start and no first
// source pass), DELSOURCE is zero with OSHPTR = SHPTR, and the next solve
// starts from the radiance truncation of the loaded radiance.
int state_load(State& s, const int32_t* shptr, const float* source, const int32_t* rshptr,
               const float* radiance, const float* fluxes, std::string& err) {
  const Plan& p = *s.plan;
  const int npts = p.npts, nst = s.nst;
  {
    int rc = check_ptr(shptr, npts, p.nlm, 0, "SHPTR", err);
    if (!rc) rc = check_ptr(rshptr, npts, p.nlm, 1, "RSHPTR", err);
    if (rc) return rc;
  }
  // the Stokes components one after another: source [nst][SHPTR(npts)],
  // radiance [nst][RSHPTR(npts)]
  const size_t ns_tot = (size_t)shptr[npts], nr_tot = (size_t)rshptr[npts];
  const size_t bytes = ((ns_tot + nr_tot) * nst + 2 * ((size_t)npts + 1)) * 4 + 1024;
  void* tmp = nullptr;
  HIPCHK(dev_alloc(&tmp, bytes));
  char* B = static_cast<char*>(tmp);
  float* d_src = reinterpret_cast<float*>(B);
  float* d_rad = d_src + ns_tot * nst;
  int32_t* d_sp = reinterpret_cast<int32_t*>(d_rad + nr_tot * nst);
  int32_t* d_rp = d_sp + npts + 1;
  int rc = AT3D_OK;
  auto run = [&]() -> int {
    HIPCHK(hipStreamSynchronize(s.stream));
    if (ns_tot) HIPCHK(hipMemcpyAsync(d_src, source, ns_tot * nst * 4, hipMemcpyHostToDevice, s.stream));
    HIPCHK(hipMemcpyAsync(d_rad, radiance, nr_tot * nst * 4, hipMemcpyHostToDevice, s.stream));
    HIPCHK(hipMemcpyAsync(d_sp, shptr, (npts + 1) * 4, hipMemcpyHostToDevice, s.stream));
    HIPCHK(hipMemcpyAsync(d_rp, rshptr, (npts + 1) * 4, hipMemcpyHostToDevice, s.stream));
    const size_t n = (size_t)s.nptsp * s.ld;
    const int nb = (int)((n + 255) / 256);
    for (int k = 0; k < nst; k++) {
      k_expand_sh<<<nb, 256, 0, s.stream>>>(d_src + ns_tot * k, d_sp, npts, s.nptsp, s.ld, s.S + s.psh * k, s.ns);
      k_expand_sh<<<nb, 256, 0, s.stream>>>(d_rad + nr_tot * k, d_rp, npts, s.nptsp, s.ld, s.R + s.psh * k, s.nr);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemsetAsync(s.DS, 0, s.psh * nst * sizeof(float), s.stream));
    HIPCHK(hipMemcpyAsync(s.nos, s.ns, npts * sizeof(int), hipMemcpyDeviceToDevice, s.stream));
    HIPCHK(hipMemcpyAsync(s.fluxes, fluxes, 2 * (size_t)npts * 4, hipMemcpyHostToDevice, s.stream));
    HIPCHK(hipStreamSynchronize(s.stream));
    return AT3D_OK;
  };
  rc = run();
  dev_free(tmp, bytes);
  if (rc) return rc;
  // the next solve remakes the radiance truncation from this radiance
  s.nr_next_fixsh = true;
  s.solcrit = 1.0f;
  s.iters = 0;
  s.initialized = true;
  return AT3D_OK;
}

// The current solution in the host format.  shptr and rshptr are always
// written; source and radiance only when given, and then their capacities
// (in floats) must hold SHPTR[npts] and RSHPTR[npts] terms.
int state_save(State& s, int32_t* shptr, int32_t* rshptr, float* source, int64_t source_cap,
               float* radiance, int64_t radiance_cap, float* fluxes, std::string& err) {
  const Plan& p = *s.plan;
  const int npts = p.npts, nst = s.nst;
  if (!s.initialized) {
    err = "no solution to save: initialise or solve first";
    return AT3D_ERR_ARG;
  }
  std::vector<int> ns(npts), nr(npts);
  HIPCHK(hipStreamSynchronize(s.stream));
  HIPCHK(hipMemcpyAsync(ns.data(), s.ns, npts * 4, hipMemcpyDeviceToHost, s.stream));
  HIPCHK(hipMemcpyAsync(nr.data(), s.nr, npts * 4, hipMemcpyDeviceToHost, s.stream));
  HIPCHK(hipStreamSynchronize(s.stream));
  shptr[0] = 0;
  rshptr[0] = 0;
  for (int i = 0; i < npts; i++) {
    shptr[i + 1] = shptr[i] + ns[i];
    rshptr[i + 1] = rshptr[i] + nr[i];
  }
  rshptr[npts + 1] = rshptr[npts];
  if (fluxes) {
    HIPCHK(hipMemcpyAsync(fluxes, s.fluxes, 2 * (size_t)npts * 4, hipMemcpyDeviceToHost, s.stream));
    HIPCHK(hipStreamSynchronize(s.stream));
  }
  // one component's dense rows into its ragged terms, in blocks of 4096 points
  auto pack_plane = [&](const float* dense, const int32_t* ptr, float* out) -> int {
    const int blk = 4096;
    std::vector<float> rows((size_t)blk * s.ld);
    for (int i0 = 0; i0 < npts; i0 += blk) {
      const int n = std::min(blk, npts - i0);
      HIPCHK(hipMemcpyAsync(rows.data(), dense + (size_t)i0 * s.ld, (size_t)n * s.ld * 4,
                            hipMemcpyDeviceToHost, s.stream));
      HIPCHK(hipStreamSynchronize(s.stream));
      for (int k = 0; k < n; k++) {
        const int i = i0 + k;
        const int c = ptr[i + 1] - ptr[i];
        for (int j = 0; j < c; j++) out[ptr[i] + j] = rows[(size_t)k * s.ld + j];
      }
    }
    return AT3D_OK;
  };
  // the Stokes components one after another, each ptr[npts] terms
  auto pack = [&](const float* dense, const int32_t* ptr, float* out, int64_t cap,
                  const char* what) -> int {
    if (!out) return AT3D_OK;
    if (cap < (int64_t)ptr[npts] * nst) {
      err = std::string(what) + ": output holds " + std::to_string(cap) + " terms, the solution " +
            std::to_string((int64_t)ptr[npts] * nst);
      return AT3D_ERR_ARG;
    }
    for (int k = 0; k < nst; k++) {
      const int rc = pack_plane(dense + s.psh * k, ptr, out + (size_t)ptr[npts] * k);
      if (rc) return rc;
    }
    return AT3D_OK;
  };
  int rc = pack(s.S, shptr, source, source_cap, "SOURCE");
  if (!rc) rc = pack(s.R, rshptr, radiance, radiance_cap, "RADIANCE");
  return rc;
}

// Mean radiance and net fluxes per point (COMPUTE_SH) into shterms
// [nshout][npts] and the net flux divergence (COMPUTE_NETFLUXDIV) into
// netfluxdiv [npts]; either may be null.
int state_sh_terms(State& s, int nshout, float* shterms, float* netfluxdiv, std::string& err) {
  const Plan& p = *s.plan;
  const int npts = p.npts;
  // (a polarized solve: the I terms, RADIANCE(1,*) of COMPUTE_SH and
  // COMPUTE_NETFLUXDIV)
  if (!s.initialized) {
    err = "no solution: initialise or solve first";
    return AT3D_ERR_ARG;
  }
  if (nshout != 4 && nshout != 5) {
    err = "nshout must be 4 or 5";
    return AT3D_ERR_ARG;
  }
  const size_t bytes = ((size_t)nshout + 1) * npts * 4;
  void* tmp = nullptr;
  HIPCHK(dev_alloc(&tmp, bytes));
  float* d_sh = static_cast<float*>(tmp);
  float* d_nfd = d_sh + (size_t)nshout * npts;
  auto run = [&]() -> int {
    k_sh_terms<<<(npts + 255) / 256, 256, 0, s.stream>>>(s.R, s.ld, s.nr, s.dirflux, s.ysun, s.ext,
                                                         s.alb, npts, p.mm, p.d.solarmu,
                                                         p.d.solaraz, nshout, d_sh, d_nfd);
    HIPCHK(hipGetLastError());
    if (shterms)
      HIPCHK(hipMemcpyAsync(shterms, d_sh, (size_t)nshout * npts * 4, hipMemcpyDeviceToHost, s.stream));
    if (netfluxdiv)
      HIPCHK(hipMemcpyAsync(netfluxdiv, d_nfd, (size_t)npts * 4, hipMemcpyDeviceToHost, s.stream));
    HIPCHK(hipStreamSynchronize(s.stream));
    return AT3D_OK;
  };
  const int rc = run();
  dev_free(tmp, bytes);
  return rc;
}

}  // namespace at3d
