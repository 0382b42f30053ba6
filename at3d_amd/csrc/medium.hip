// The medium of a solve set up on the device (TRANSFER_PA_TO_GRID /
// PREPARE_PROP, shdomsub1.f:19-112 and shdomsub2.f:479-608; the source and
// single-scatter constants; the two-stream start, INIT_RADIANCE with EDDRTF,
// shdomsub2.f:614-993; the delta-M scaled property-grid extinction of
// DIRECT_BEAM_PROP): one thread per grid point, column or property-grid
// point, running the per-point functions the host plan runs
// (medium_point.h), so a state gets the host's values without the host's
// per-point work.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "medium_point.h"

namespace at3d {

// Property transfer and delta-M of every point: species-major optics, the
// total extinction and the largest albedo (as float bits; albedos are >= 0).
__global__ __launch_bounds__(256) void k_medium_optics(MediumArgs m) {
#pragma clang fp contract(off)
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= m.npts) return;
  const size_t np = m.npts;
  const float X = m.gridpos[(size_t)ip * 3], Y = m.gridpos[(size_t)ip * 3 + 1], Z = m.gridpos[(size_t)ip * 3 + 2];
  float total = 0.f;
  for (int ipa = 0; ipa < m.npart; ipa++) {
    float e, a;
    prop_point(m.g, ipa, X, Y, Z, e, a, m.iphase + (ipa * np + ip) * m.t.npw, m.wt + (ipa * np + ip) * m.t.npw);
    m.extinct[ipa * np + ip] = e;
    m.albedo[ipa * np + ip] = a;
    total = total + e;
  }
  if (m.deltam) {
    float s = 0.f;
    for (int ipa = 0; ipa < m.npart; ipa++) s += m.extinct[ipa * np + ip];
    const float t = total - s;
    total = t < 0.f ? 0.f : t;
    for (int ipa = 0; ipa < m.npart; ipa++) {
      const float f = prepare_f(m.t, m.iphase + (ipa * np + ip) * m.t.npw, m.wt + (ipa * np + ip) * m.t.npw);
      float e = m.extinct[ipa * np + ip], a = m.albedo[ipa * np + ip];
      deltam_scale(f, e, a);
      m.extinct[ipa * np + ip] = e;
      m.albedo[ipa * np + ip] = a;
      total = total + e;
    }
  }
  m.total_ext[ip] = total;
  unsigned int amax = 0;
  for (int ipa = 0; ipa < m.npart; ipa++) amax = max(amax, __float_as_uint(m.albedo[ipa * np + ip]));
  atomicMax(m.albmax, amax);
}

// Source constants, the source class (highest order with a nonzero
// albedo-weighted phase term), the TMS terms and the number of exact
// single-scatter pairs of every point.
__global__ __launch_bounds__(256) void k_medium_constants(MediumArgs m) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= m.npts) return;
  const size_t np = m.npts;
  const int ml1 = m.t.ml + 1;
  float* chi = m.chi + (size_t)ip * ml1;
  float alb;
  point_source_constants(m.t, m.npart, m.deltam, m.total_ext[ip], m.extinct + ip, m.albedo + ip,
                         m.iphase + (size_t)ip * m.t.npw, m.wt + (size_t)ip * m.t.npw, np, chi, alb,
                         m.tc + (size_t)ip * ml1);
  m.alb[ip] = alb;
  int c = 0;
  for (int l = m.t.ml; l > 0; l--)
    if (alb * chi[l] != 0.f) {
      c = l;
      break;
    }
  m.cls[ip] = c;
  int n = 0;
  if (m.deltam)
    n = point_render_constants(m.t, m.npart, m.total_ext[ip], m.extinct + ip, m.albedo + ip,
                               m.iphase + (size_t)ip * m.t.npw, m.wt + (size_t)ip * m.t.npw, np, m.secmu0,
                               m.tms + (size_t)ip * ml1, nullptr, nullptr);
  m.ss_cnt[ip] = n;
}

// The exact single-scatter pairs of every point at its CSR offset.
__global__ __launch_bounds__(256) void k_medium_pairs(MediumArgs m) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= m.npts || !m.deltam) return;
  const int o = m.ss_off[ip];
  point_render_constants(m.t, m.npart, m.total_ext[ip], m.extinct + ip, m.albedo + ip,
                         m.iphase + (size_t)ip * m.t.npw, m.wt + (size_t)ip * m.t.npw, m.npts, m.secmu0, nullptr,
                         m.ss_iph + o, m.ss_coef + o);
}

// Exclusive prefix sum of the pair counts (npts + 1 offsets): sums of
// 1024-count blocks, their prefix in one workgroup, then each block's own.
constexpr int SCAN_BLOCK = 1024;

__global__ __launch_bounds__(256) void k_scan_blocks(const int32_t* __restrict__ cnt, int n,
                                                     int32_t* __restrict__ block_sum) {
  __shared__ int32_t part[256];
  const int b = blockIdx.x, t = threadIdx.x;
  int32_t acc = 0;
  for (int k = t; k < SCAN_BLOCK; k += 256) {
    const int i = b * SCAN_BLOCK + k;
    if (i < n) acc += cnt[i];
  }
  part[t] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) part[t] += part[t + s];
    __syncthreads();
  }
  if (t == 0) block_sum[b] = part[0];
}

__global__ __launch_bounds__(64) void k_scan_tops(int32_t* __restrict__ block_sum, int nblocks,
                                                  int32_t* __restrict__ total) {
  if (threadIdx.x != 0) return;
  int32_t acc = 0;
  for (int b = 0; b < nblocks; b++) {
    const int32_t v = block_sum[b];
    block_sum[b] = acc;
    acc += v;
  }
  total[0] = acc;
}

__global__ __launch_bounds__(64) void k_scan_fill(const int32_t* __restrict__ cnt, int n,
                                                  const int32_t* __restrict__ block_sum,
                                                  const int32_t* __restrict__ total, int32_t* __restrict__ off) {
  // one lane per block walks its 1024 counts (the counts are small and the
  // pass runs once per state)
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int nblocks = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
  if (b >= nblocks) return;
  int32_t acc = block_sum[b];
  for (int k = 0; k < SCAN_BLOCK; k++) {
    const int i = b * SCAN_BLOCK + k;
    if (i >= n) break;
    off[i] = acc;
    acc += cnt[i];
  }
  if (b == nblocks - 1) off[n] = total[0];
}

// The two-stream start of every column into the radiance's first terms.
__global__ __launch_bounds__(64) void k_medium_eddington(MediumArgs m) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= m.ncol) return;
  const int nz = m.nz, nlayer = nz - 1;
  const size_t nw = 2 * (size_t)nlayer + 3;
  double* wk = m.work + (size_t)col * 4 * nw;
  float* fw = m.fwork + (size_t)col * (3 * (size_t)nz + 3 * (size_t)(nlayer + 1));
  float* od = fw;
  float* alb = fw + nz;
  float* asym = fw + 2 * nz;
  float* fl = fw + 3 * nz;
  eddington_layers(m.t, m.npart, m.deltam, nz, m.zgrid, m.total_ext, m.extinct, m.albedo, m.iphase, m.wt, m.npts,
                   col, od, alb, asym);
  if (!eddrtf_column(nlayer, od, alb, asym, m.solarflux, m.solarmu, m.gndemis, m.skyflux_weight, fl, wk, wk + nw,
                     wk + 2 * nw, wk + 3 * nw)) {
    atomicOr(m.err, 1);
    return;
  }
  column_radiance(nz, col, fl, m.radiance0);
}

// The delta-M scaled extinction of the property grid that the direct beam
// integrates (DIRECT_BEAM_PROP's set-up).
__global__ __launch_bounds__(256) void k_medium_beam_ext(MediumArgs m) {
#pragma clang fp contract(off)
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  const int npg = m.g.npx * m.g.npy * m.g.npz;
  if (ip >= npg) return;
  float acc = 0.f;
  for (int ipa = 0; ipa < m.npart; ipa++) {
    double ext = m.g.extinctp[(size_t)ipa * npg + ip];
    const double alb = m.g.albedop[(size_t)ipa * npg + ip];
    if (m.deltam) {
      const int L = m.t.ml + 1;
      double F = 0.0;
      for (int q = 0; q < m.g.nmicro; q++) {
        const size_t pq = ((size_t)ipa * npg + ip) * m.g.nmicro + q;
        const int iph = m.g.iphasep[pq];
        float term = m.g.phasewtp[pq] * m.legenp[(size_t)(iph - 1) * (m.nlegp + 1) + L];
        term = term / (float)(2 * L + 1);
        F = F + (double)term;
      }
      ext = (1.0 - alb * F) * ext;
    }
    acc = (float)((double)acc + ext);
  }
  m.beam_ext[ip] = acc;
}

void launch_medium_points(hipStream_t st, const MediumArgs& m, int32_t* block_sum, int32_t* total) {
  const int nb = (m.npts + 255) / 256;
  k_medium_optics<<<nb, 256, 0, st>>>(m);
  k_medium_constants<<<nb, 256, 0, st>>>(m);
  const int nblocks = (m.npts + SCAN_BLOCK - 1) / SCAN_BLOCK;
  k_scan_blocks<<<nblocks, 256, 0, st>>>(m.ss_cnt, m.npts, block_sum);
  k_scan_tops<<<1, 64, 0, st>>>(block_sum, nblocks, total);
  k_scan_fill<<<(nblocks + 63) / 64, 64, 0, st>>>(m.ss_cnt, m.npts, block_sum, total, m.ss_off);
}

void launch_medium_pairs(hipStream_t st, const MediumArgs& m) {
  k_medium_pairs<<<(m.npts + 255) / 256, 256, 0, st>>>(m);
}

void launch_medium_start(hipStream_t st, const MediumArgs& m) {
  const int npg = m.g.npx * m.g.npy * m.g.npz;
  k_medium_beam_ext<<<(npg + 255) / 256, 256, 0, st>>>(m);
  k_medium_eddington<<<(m.ncol + 63) / 64, 64, 0, st>>>(m);
}

}  // namespace at3d
