// The direct solar beam of a solve on the device: one thread per grid point
// traces its slant path to the domain top (DIRECT_BEAM_PROP
// shdomsub1.f:4097-4334, as the host plan's trace), then the constants that
// scale with the direct flux are scaled by it.
#include <hip/hip_runtime.h>
#include <math.h>

#include "direct_beam.h"
#include "kernels.h"

namespace at3d {

// DIRFLUX = SOLARFLUX exp(-path) and the source pass's F0 = DIRFLUX / |mu0|.
__global__ __launch_bounds__(256) void k_direct_beam(DirectBeamGeom b, const float* __restrict__ gridpos,
                                                     int npts, float secmu0, float* __restrict__ dirflux,
                                                     float* __restrict__ f0, int* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  double path = 0.0;
  float df = 0.f;
  if (direct_beam_trace(b, gridpos[(size_t)i * 3], gridpos[(size_t)i * 3 + 1], gridpos[(size_t)i * 3 + 2], path,
                        NoBeamRecord{}))
    df = (float)(b.solarflux * exp(-path));
  else
    err[0] = 1;
  dirflux[i] = df;
  f0[i] = df * secmu0;
}

// The single-scatter constants held per unit direct flux (TMS terms and the
// per-point phase coefficients) times the point's direct flux.
__global__ __launch_bounds__(256) void k_direct_scale(const float* __restrict__ dirflux, int npts, int ml1,
                                                      float* __restrict__ tms, const int32_t* __restrict__ ss_off,
                                                      float* __restrict__ ss_coef) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  const float d = dirflux[i];
  for (int l = 0; l < ml1; l++) tms[(size_t)i * ml1 + l] *= d;
  if (ss_coef)
    for (int k = ss_off[i]; k < ss_off[i + 1]; k++) ss_coef[k] *= d;
}

void launch_direct_beam(hipStream_t st, const DirectBeamGeom& b, const float* gridpos, int npts, float secmu0,
                        float* dirflux, float* f0, int* err) {
  if (npts <= 0) return;
  k_direct_beam<<<(npts + 255) / 256, 256, 0, st>>>(b, gridpos, npts, secmu0, dirflux, f0, err);
}

void launch_direct_scale(hipStream_t st, const float* dirflux, int npts, int ml1, float* tms,
                         const int32_t* ss_off, float* ss_coef) {
  if (npts <= 0) return;
  k_direct_scale<<<(npts + 255) / 256, 256, 0, st>>>(dirflux, npts, ml1, tms, ss_off, ss_coef);
}

}  // namespace at3d
