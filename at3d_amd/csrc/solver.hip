// Device-resident SHDOM source-function iteration for gfx950.
//
// One iteration (SOLUTION_ITERATIONS loop body, shdomsub1.f:674-797):
//   radiance truncation  -> k_radiance_truncation   (RADIANCE_TRUNCATION :1615)
//   SH -> DO             -> k_sh_to_do   FP32 MFMA GEMM [nang x nlm] x [nlm x npts]
//                           with the >=0 clamp and the *extinction fused
//                           (SH_TO_DO_UNPOL :2928, PATH_INTEGRATION :2020-2037)
//   downward sweep       -> k_sweep      one workgroup per ordinate, grid points
//                           in dependency levels (BACK_INT_GRID3D_UNPOL :3695)
//   Lambertian surface   -> k_bottom_bc  (FIXED_LAMBERTIAN_BOUNDARY :2438)
//   upward sweep         -> k_sweep
//   DO -> SH             -> k_do_to_sh   FP32 MFMA GEMM [npts x nang] x [nang x nlm]
//                           masked by the radiance truncation (DO_TO_SH_UNPOL :3164)
//   source function      -> k_compute_source  one wavefront per point, FP64
//                           block partials of the acceleration dot products
//                           (COMPUTE_SOURCE :967)
//   criterion            -> k_reduce_dots + host CALC_ACCEL_SOLCRIT (shdom_nompi.f:317)
//   acceleration         -> k_accelerate (ACCELERATE_SOLUTION :1807)
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "host_plan.h"
#include "shdom_geom.h"
#include "solver_state.h"

namespace at3d {

typedef float float4_ __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// FP32 MFMA GEMM tiles.  Workgroup tile 64x64, K step 16, 256 threads = 4
// waves in a 2x2 arrangement, each wave 32x32 = 2x2 blocks of
// v_mfma_f32_16x16x4_f32.  Operand lane layout of that instruction:
//   A[i][k] in lane i + 16*k,  B[k][j] in lane j + 16*k,
//   D[4*(lane/16)+v][lane%16] in accumulator element v.
// ---------------------------------------------------------------------------
constexpr int TM = 64, TN = 64, TK = 16, LDS_PAD = 4;

// DOsrc[k][p] = max(0, sum_J Y[k][J] * S[p][J]) * ext[p]
// Y: [Mpad][ldk] (ordinates), S: [Npad][ldk] (points), K = ldk (multiple of 16).
__global__ __launch_bounds__(256) void k_sh_to_do(const float* __restrict__ Y,
                                                  const float* __restrict__ S,
                                                  const float* __restrict__ ext,
                                                  float* __restrict__ out, int ldk, int ldo,
                                                  int nang, int npts) {
  __shared__ float sA[TK][TM + LDS_PAD];
  __shared__ float sB[TK][TN + LDS_PAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  float4_ acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = (float4_){0.f, 0.f, 0.f, 0.f};
  const int lr = tid >> 2, lc = (tid & 3) * 4;  // tile row, k offset for loads
  for (int k0 = 0; k0 < ldk; k0 += TK) {
    const float4 a = *reinterpret_cast<const float4*>(Y + (size_t)(m0 + lr) * ldk + k0 + lc);
    const float4 b = *reinterpret_cast<const float4*>(S + (size_t)(n0 + lr) * ldk + k0 + lc);
    __syncthreads();
    sA[lc + 0][lr] = a.x;
    sA[lc + 1][lr] = a.y;
    sA[lc + 2][lr] = a.z;
    sA[lc + 3][lr] = a.w;
    sB[lc + 0][lr] = b.x;
    sB[lc + 1][lr] = b.y;
    sB[lc + 2][lr] = b.z;
    sB[lc + 3][lr] = b.w;
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TK; kk += 4) {
      const int kr = kk + (lane >> 4);
      float av[2], bv[2];
      for (int i = 0; i < 2; i++) av[i] = sA[kr][wm * 32 + i * 16 + (lane & 15)];
      for (int j = 0; j < 2; j++) bv[j] = sB[kr][wn * 32 + j * 16 + (lane & 15)];
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], bv[j], acc[i][j], 0, 0, 0);
    }
  }
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) {
      const int p = n0 + wn * 32 + j * 16 + (lane & 15);
      if (p >= npts) continue;
      const float e = ext[p];
      for (int v = 0; v < 4; v++) {
        const int k = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + v;
        if (k < nang) out[(size_t)k * ldo + p] = fmaxf(acc[i][j][v], 0.0f) * e;
      }
    }
}

// R[p][J] = (J < nr[p]) ? sum_k D[k][p] * W[k][J] : 0
// D: [Kpad][ldd] (ordinate-major DO radiance), W: [Kpad][ldw]; K = kpad (multiple of 16).
__global__ __launch_bounds__(256) void k_do_to_sh(const float* __restrict__ D,
                                                  const float* __restrict__ W,
                                                  const int* __restrict__ nr,
                                                  float* __restrict__ R, int kpad, int ldd,
                                                  int ldw, int npts) {
  __shared__ float sA[TK][TM + LDS_PAD];
  __shared__ float sB[TK][TN + LDS_PAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.x * TM, n0 = blockIdx.y * TN;
  float4_ acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = (float4_){0.f, 0.f, 0.f, 0.f};
  const int lk = tid >> 4, lc = (tid & 15) * 4;
  for (int k0 = 0; k0 < kpad; k0 += TK) {
    const float4 a = *reinterpret_cast<const float4*>(D + (size_t)(k0 + lk) * ldd + m0 + lc);
    const float4 b = *reinterpret_cast<const float4*>(W + (size_t)(k0 + lk) * ldw + n0 + lc);
    __syncthreads();
    *reinterpret_cast<float4*>(&sA[lk][lc]) = a;
    *reinterpret_cast<float4*>(&sB[lk][lc]) = b;
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TK; kk += 4) {
      const int kr = kk + (lane >> 4);
      float av[2], bv[2];
      for (int i = 0; i < 2; i++) av[i] = sA[kr][wm * 32 + i * 16 + (lane & 15)];
      for (int j = 0; j < 2; j++) bv[j] = sB[kr][wn * 32 + j * 16 + (lane & 15)];
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], bv[j], acc[i][j], 0, 0, 0);
    }
  }
  for (int i = 0; i < 2; i++)
    for (int v = 0; v < 4; v++) {
      const int p = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + v;
      if (p >= npts) continue;
      const int n = nr[p];
      for (int j = 0; j < 2; j++) {
        const int J = n0 + wn * 32 + j * 16 + (lane & 15);
        R[(size_t)p * ldw + J] = (J < n) ? acc[i][j][v] : 0.0f;
      }
    }
}

// ---------------------------------------------------------------------------
// Discrete-ordinate sweep.  One workgroup integrates one ordinate over the
// whole grid; points are processed level by level (every upwind face point of
// a level-L point has a level < L), with a workgroup barrier between levels.
// The characteristic itself is the reference's: trace back from the point
// through cells until a face whose four corner radiances are already known in
// the reference's sweep order, integrating each cell crossed.
// ---------------------------------------------------------------------------
struct SweepArgs {
  GridTopo g;
  const OrdGeom* ordg;        // [nang]
  const int32_t* ord_oct;     // [nang]
  const float* skyrad;        // [nang] top boundary radiance (down ordinates)
  const float* bcrad_bot;     // [nbot] bottom boundary radiance (up ordinates)
  const int32_t* bctop;       // [ntop]
  const int32_t* bcbot;       // [nbot]
  int ntop, nbot;
  const int8_t* preset;       // [npts] bit0 top point, bit1 bottom point
  const int32_t* rank;        // [8][npts]
  const int32_t* level;       // [8][npts]
  const int32_t* lvl_pts;     // concatenated over octants
  const int32_t* lvl_cell;
  const int32_t* lvl_off;     // concatenated offsets
  const int32_t* oct_lvl_base;  // [8] start of each octant's offsets in lvl_off
  const int32_t* oct_pts_base;  // [8] start of each octant's points in lvl_pts
  const int32_t* oct_nlev;      // [8]
  const float* ext;           // [npts]
  const float* srcext;        // [nangp][ldo]
  float* rad;                 // [nangp][ldo]
  int ldo, npts;
  int* errflag;
};

__global__ __launch_bounds__(1024) void k_sweep(SweepArgs a, int kbase) {
  const int k = kbase + blockIdx.x;
  const OrdGeom og = a.ordg[k];
  const int o = og.ioct;
  const bool up = og.bitz != 0;
  const int presetbit = up ? 2 : 1;
  float* __restrict__ rad = a.rad + (size_t)k * a.ldo;
  const float* __restrict__ se = a.srcext + (size_t)k * a.ldo;
  const float* __restrict__ ext = a.ext;
  if (!up) {
    const float v = a.skyrad[k];
    for (int i = threadIdx.x; i < a.ntop; i += blockDim.x) rad[a.bctop[i]] = v;
  } else {
    for (int i = threadIdx.x; i < a.nbot; i += blockDim.x) rad[a.bcbot[i]] = a.bcrad_bot[i];
  }
  __syncthreads();
  const int32_t* rank = a.rank + (size_t)o * a.npts;
  const int32_t* level = a.level + (size_t)o * a.npts;
  const int32_t* off = a.lvl_off + a.oct_lvl_base[o];
  const int32_t* pts = a.lvl_pts + a.oct_pts_base[o];
  const int32_t* cells = a.lvl_cell + a.oct_pts_base[o];
  const int nlev = a.oct_nlev[o];
  for (int L = 1; L <= nlev; L++) {
    const int beg = off[L - 1], end = off[L];
    for (int i = beg + (int)threadIdx.x; i < end; i += blockDim.x) {
      const int pt = pts[i];
      const int rp = rank[pt];
      int icell = cells[i];
      double xe = a.g.gridpos[pt * 3], ye = a.g.gridpos[pt * 3 + 1], ze = a.g.gridpos[pt * 3 + 2];
      double ext1 = ext[pt], srcext1 = se[pt];
      double radv = 0.0, transmit = 1.0;
      bool done = false;
      for (int step = 0; step <= a.g.ncells; step++) {
        TraceStep t;
        if (!trace_step(a.g, og, icell, xe, ye, ze, t)) {
          atomicOr(a.errflag, 1);
          break;
        }
        const double f1 = (1 - t.u) * (1 - t.v), f2 = (1 - t.u) * t.v;
        const double f3 = t.u * (1 - t.v), f4 = t.u * t.v;
        const double ext0 = f1 * ext[t.i[0]] + f2 * ext[t.i[1]] + f3 * ext[t.i[2]] + f4 * ext[t.i[3]];
        const double srcext0 = f1 * se[t.i[0]] + f2 * se[t.i[1]] + f3 * se[t.i[2]] + f4 * se[t.i[3]];
        double src, absc, trc;
        cell_integral(ext0, ext1, srcext0, srcext1, t.so, src, absc, trc);
        radv = radv + transmit * src * absc;
        transmit = transmit * trc;
        bool valid = true;
        for (int c = 0; c < 4; c++) {
          const int q = t.i[c];
          valid = valid && ((a.preset[q] & presetbit) || rank[q] < rp);
        }
        if (t.next < 0 || valid) {
          if (valid) {
            for (int c = 0; c < 4; c++) {
              const int q = t.i[c];
              if (!(a.preset[q] & presetbit) && level[q] >= L) atomicOr(a.errflag, 4);
            }
            const double rad0 = f1 * rad[t.i[0]] + f2 * rad[t.i[1]] + f3 * rad[t.i[2]] + f4 * rad[t.i[3]];
            radv = radv + transmit * rad0;
          } else {
            atomicOr(a.errflag, 2);
          }
          done = true;
          break;
        }
        ext1 = ext0;
        srcext1 = srcext0;
        icell = t.next;
      }
      if (!done) atomicOr(a.errflag, 8);
      rad[pt] = (float)radv;
    }
    __syncthreads();
  }
}

// Lambertian surface radiance from the downwelling hemispheric flux
// (FIXED_LAMBERTIAN_BOUNDARY shdomsub1.f:2438-2480; flux sum :2118-2135).
__global__ void k_bottom_bc(const float* __restrict__ rad, int ldo, const int32_t* __restrict__ bcbot,
                            int nbot, const float* __restrict__ fluxw, int ndown,
                            const float* __restrict__ dirflux, float albpi,
                            float* __restrict__ bcrad) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbot) return;
  const int p = bcbot[b];
  float f = 0.f;
  for (int k = 0; k < ndown; k++) f = f + fluxw[k] * rad[(size_t)k * ldo + p];
  bcrad[b] = albpi * (dirflux[p] + f);
}

// Hemispheric fluxes at every point (PATH_INTEGRATION :2118-2135).
__global__ void k_fluxes(const float* __restrict__ rad, int ldo, const float* __restrict__ fluxw,
                         int ndown, int nang, int npts, float* __restrict__ fluxes) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  float dn = 0.f, upf = 0.f;
  for (int k = 0; k < ndown; k++) dn = dn + fluxw[k] * rad[(size_t)k * ldo + p];
  for (int k = ndown; k < nang; k++) upf = upf + fluxw[k] * rad[(size_t)k * ldo + p];
  fluxes[p] = dn;
  fluxes[npts + p] = upf;
}

__device__ inline int nterms_for_l(int l, int mm) {
  return (l <= mm) ? l * (l + 1) + l + 1 : (2 * mm + 1) * l - mm * mm + mm + 1;
}

// RADIANCE_TRUNCATION (shdomsub1.f:1615-1803).
__global__ void k_radiance_truncation(const float* __restrict__ R, int ld, const float* __restrict__ tc,
                                      const int* __restrict__ ns, const int* __restrict__ lofj,
                                      int* __restrict__ nr, int npts, int ml, int mm, int nlm,
                                      float radmin, int fixsh, int highorder) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  if (fixsh) {
    int n = max(4, ns[p]);
    if (highorder) n = nterms_for_l(ml, mm);
    nr[p] = n;
    return;
  }
  const float i0 = R[(size_t)p * ld];
  int lr = 1;
  for (int l = 1; l <= ml; l++)
    if (tc[(size_t)p * (ml + 1) + l] * i0 > radmin) lr = l;
  const int n0 = max(1, ns[p]);
  const int ls = lofj[n0 - 1];
  lr = min(lr, ls + ml / 8 + 2);
  if (highorder) lr = ml;
  nr[p] = nterms_for_l(lr, mm);
}

// COMPUTE_SOURCE for NSTOKES=1, one wavefront per grid point (shdomsub1.f:967-1611).
// first: initialisation call (no dot products or DELSOURCE).
struct SourceArgs {
  const float* R;
  const int* nr;
  const float* alb;
  const float* chi;   // [npts][ml+1]
  const float* f0;    // DIRFLUX * SECMU0
  const float* ysun;  // [nlm]
  const int* lofj;    // [nlm]
  float* S;
  float* DS;
  int* ns;
  int* nos;
  double* partial;    // [gridDim][4]
  int ld, npts, ml, mm, nlm;
  float shacc;
  int first, accel, fixsh;
};

__global__ __launch_bounds__(256) void k_compute_source(SourceArgs a) {
  __shared__ double red[4][4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int p = blockIdx.x * 4 + w;
  double dot = 0, dold = 0, dnew = 0, jn = 0;
  if (p < a.npts) {
    const float alb = a.alb[p], f0 = a.f0[p];
    const int nrp = a.nr[p], ns_old = a.ns[p], nos_old = a.nos[p];
    const float* Rp = a.R + (size_t)p * a.ld;
    float* Sp = a.S + (size_t)p * a.ld;
    float* DSp = a.DS + (size_t)p * a.ld;
    const float* chip = a.chi + (size_t)p * (a.ml + 1);
    int js = -1;
    const int nmax = (a.nlm + 63) / 64;
    float st[8];
    for (int it = 0; it < nmax; it++) {
      const int J = lane + 64 * it;
      float v = 0.f;
      if (J < a.nlm) {
        const float c = chip[a.lofj[J]];
        v = f0 * alb * c * a.ysun[J];
        if (J < nrp) v = v + alb * c * Rp[J];
        if (fabsf(v) > a.shacc) js = J;
      }
      st[it] = v;
    }
    if (!a.first) {
      if (a.accel) {
        const int nsd = min(ns_old, nos_old);
        for (int it = 0; it < nmax; it++) {
          const int J = lane + 64 * it;
          if (J < nsd) {
            const double so = Sp[J], d = DSp[J];
            const double dn = (double)(st[it] - Sp[J]);
            dot += dn * d;
            dold += d * d;
            dnew += dn * dn;
            jn += so * so;
          }
        }
      } else {
        for (int it = 0; it < nmax; it++) {
          const int J = lane + 64 * it;
          if (J < ns_old) {
            const double so = Sp[J];
            const double dn = (double)(st[it] - Sp[J]);
            dnew += dn * dn;
            jn += so * so;
          }
        }
      }
      if (a.accel) {
        for (int it = 0; it < nmax; it++) {
          const int J = lane + 64 * it;
          if (J < a.ld) DSp[J] = (J < ns_old) ? st[it] - Sp[J] : 0.f;
        }
        if (lane == 0) a.nos[p] = ns_old;
      }
    }
    // new truncation: last term above SHACC rounded up to a full l
    int nsn;
    if (a.fixsh) {
      nsn = ns_old;
    } else {
      for (int off = 32; off > 0; off >>= 1) js = max(js, __shfl_xor(js, off));
      nsn = (js < 0) ? 0 : nterms_for_l(a.lofj[js], a.mm);
    }
    for (int it = 0; it < nmax; it++) {
      const int J = lane + 64 * it;
      if (J < a.ld) Sp[J] = (J < nsn && J < a.nlm) ? st[it] : 0.f;
    }
    if (lane == 0) a.ns[p] = nsn;
    if (a.first && a.accel) {
      for (int it = 0; it < nmax; it++) {
        const int J = lane + 64 * it;
        if (J < a.ld) DSp[J] = 0.f;
      }
      if (lane == 0) a.nos[p] = nsn;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    dot += __shfl_xor(dot, off);
    dold += __shfl_xor(dold, off);
    dnew += __shfl_xor(dnew, off);
    jn += __shfl_xor(jn, off);
  }
  if (lane == 0) {
    red[w][0] = dot;
    red[w][1] = dold;
    red[w][2] = dnew;
    red[w][3] = jn;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int c = threadIdx.x;
    a.partial[(size_t)blockIdx.x * 4 + c] = red[0][c] + red[1][c] + red[2][c] + red[3][c];
  }
}

// Deterministic reduction of the block partials (fixed order).
__global__ __launch_bounds__(256) void k_reduce_dots(const double* __restrict__ partial, int nblocks,
                                                     double* __restrict__ out) {
  __shared__ double sh[256][4];
  double s[4] = {0, 0, 0, 0};
  for (int b = threadIdx.x; b < nblocks; b += 256)
    for (int c = 0; c < 4; c++) s[c] += partial[(size_t)b * 4 + c];
  for (int c = 0; c < 4; c++) sh[threadIdx.x][c] = s[c];
  __syncthreads();
  for (int stride = 128; stride > 0; stride >>= 1) {
    if ((int)threadIdx.x < stride)
      for (int c = 0; c < 4; c++) sh[threadIdx.x][c] += sh[threadIdx.x + stride][c];
    __syncthreads();
  }
  if (threadIdx.x < 4) out[threadIdx.x] = sh[0][threadIdx.x];
}

// ACCELERATE_SOLUTION (shdomsub1.f:1807-1831).
__global__ void k_accelerate(float* __restrict__ S, const float* __restrict__ DS,
                             const int* __restrict__ ns, const int* __restrict__ nos, int ld,
                             int npts, float a) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t p = idx / ld;
  const int J = (int)(idx % ld);
  if (p >= (size_t)npts) return;
  if (J < min(ns[p], nos[p])) S[idx] = S[idx] + a * DS[idx];
}

__global__ void k_fill_f32(float* __restrict__ x, size_t n, float v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
#define HIPCHK(x)                                                              \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      err = std::string(#x) + ": " + hipGetErrorString(e_);                    \
      return AT3D_ERR_DEVICE;                                                  \
    }                                                                          \
  } while (0)

static size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

template <typename T>
static int upload(State& s, T*& dst, const T* src, size_t n, std::string& err) {
  HIPCHK(hipMalloc((void**)&dst, std::max<size_t>(n, 1) * sizeof(T)));
  s.allocs.push_back(dst);
  if (n) HIPCHK(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return AT3D_OK;
}

template <typename T>
static int alloc0(State& s, T*& dst, size_t n, std::string& err) {
  HIPCHK(hipMalloc((void**)&dst, std::max<size_t>(n, 1) * sizeof(T)));
  s.allocs.push_back(dst);
  HIPCHK(hipMemset(dst, 0, std::max<size_t>(n, 1) * sizeof(T)));
  return AT3D_OK;
}

#define TRY(x)              \
  do {                      \
    int rc_ = (x);          \
    if (rc_) return rc_;    \
  } while (0)

int state_create(const Plan& p, State& s, std::string& err) {
  s.plan = &p;
  s.npts = p.npts;
  s.nlm = p.nlm;
  s.nang = p.nang;
  s.ld = (int)round_up(p.nlm, 64);
  s.nptsp = (int)round_up(p.npts, 64);
  s.nangp = (int)round_up(p.nang, 64);
  s.ndown = 0;
  for (const Ordinate& o : p.ords) s.ndown += o.down ? 1 : 0;
  if (p.nlm > 512) {
    err = "NLM > 512 is not supported by the source kernel";
    return AT3D_ERR_ARG;
  }
  HIPCHK(hipStreamCreate(&s.stream));
  // transform matrices, zero padded
  std::vector<float> Y((size_t)s.nangp * s.ld, 0.f), W((size_t)s.nangp * s.ld, 0.f);
  for (int k = 0; k < p.nang; k++)
    for (int j = 0; j < p.nlm; j++) {
      Y[(size_t)k * s.ld + j] = p.ymat[(size_t)k * p.nlm + j];
      W[(size_t)k * s.ld + j] = p.wmat[(size_t)k * p.nlm + j];
    }
  TRY(upload(s, s.Y, Y.data(), Y.size(), err));
  TRY(upload(s, s.W, W.data(), W.size(), err));
  std::vector<float> ysun(s.ld, 0.f);
  std::copy(p.ylmsun.begin(), p.ylmsun.end(), ysun.begin());
  TRY(upload(s, s.ysun, ysun.data(), ysun.size(), err));
  TRY(upload(s, s.lofj, p.lofj.data(), p.lofj.size(), err));
  // per-point optics (padded to nptsp)
  std::vector<float> ext(s.nptsp, 0.f);
  std::copy(p.total_ext.begin(), p.total_ext.end(), ext.begin());
  TRY(upload(s, s.ext, ext.data(), ext.size(), err));
  TRY(upload(s, s.alb, p.alb.data(), p.alb.size(), err));
  TRY(upload(s, s.chi, p.chi.data(), p.chi.size(), err));
  TRY(upload(s, s.tc, p.tc.data(), p.tc.size(), err));
  TRY(upload(s, s.dirflux, p.dirflux.data(), p.dirflux.size(), err));
  std::vector<float> f0(p.npts);
  const float secmu0 = 1.0f / fabsf(p.d.solarmu);
  for (int i = 0; i < p.npts; i++) f0[i] = p.dirflux[i] * secmu0;
  TRY(upload(s, s.f0, f0.data(), f0.size(), err));
  // big arrays
  const size_t nsh = (size_t)s.nptsp * s.ld, ndo = (size_t)s.nangp * s.nptsp;
  TRY(alloc0(s, s.S, nsh, err));
  TRY(alloc0(s, s.R, nsh, err));
  TRY(alloc0(s, s.DS, nsh, err));
  TRY(alloc0(s, s.dosrc, ndo, err));
  TRY(alloc0(s, s.dorad, ndo, err));
  TRY(alloc0(s, s.ns, (size_t)p.npts, err));
  TRY(alloc0(s, s.nr, (size_t)p.npts, err));
  TRY(alloc0(s, s.nos, (size_t)p.npts, err));
  TRY(alloc0(s, s.fluxes, (size_t)2 * p.npts, err));
  s.nsrc_blocks = (p.npts + 3) / 4;
  TRY(alloc0(s, s.partial, (size_t)s.nsrc_blocks * 4, err));
  TRY(alloc0(s, s.dots, (size_t)4, err));
  TRY(alloc0(s, s.errflag, (size_t)1, err));
  // sweep geometry
  TRY(upload(s, s.gridpos, p.gridpos.data(), p.gridpos.size(), err));
  TRY(upload(s, s.gridptr, p.gridptr.data(), p.gridptr.size(), err));
  TRY(upload(s, s.neighptr, p.neighptr.data(), p.neighptr.size(), err));
  TRY(upload(s, s.cellflags, p.cellflags.data(), p.cellflags.size(), err));
  std::vector<OrdGeom> og(p.nang);
  std::vector<int32_t> oo(p.nang);
  std::vector<float> sky(p.nang, 0.f), fluxw(p.nang);
  for (int k = 0; k < p.nang; k++) {
    og[k] = p.ords[k].g;
    oo[k] = p.ords[k].g.ioct;
    sky[k] = p.ords[k].down ? p.d.skyrad : 0.f;
    fluxw[k] = fabsf(p.ords[k].mu) * p.ords[k].wtdo;
  }
  // ordinates must be down first (the reference's ordering of MU)
  for (int k = 0; k < p.nang; k++)
    if (p.ords[k].down != (k < s.ndown)) {
      err = "internal: downwelling ordinates are not first";
      return AT3D_ERR_ARG;
    }
  TRY(upload(s, s.ordg, og.data(), og.size(), err));
  TRY(upload(s, s.ordoct, oo.data(), oo.size(), err));
  TRY(upload(s, s.skyrad, sky.data(), sky.size(), err));
  TRY(upload(s, s.fluxw, fluxw.data(), fluxw.size(), err));
  TRY(upload(s, s.bctop, p.bctop.data(), p.bctop.size(), err));
  TRY(upload(s, s.bcbot, p.bcbot.data(), p.bcbot.size(), err));
  TRY(alloc0(s, s.bcrad_bot, p.bcbot.size(), err));
  std::vector<int8_t> preset(p.npts, 0);
  for (int i : p.bctop) preset[i] |= 1;
  for (int i : p.bcbot) preset[i] |= 2;
  TRY(upload(s, s.preset, preset.data(), preset.size(), err));
  std::vector<int32_t> rank, level, lp, lc, lo, lvl_base(8), pts_base(8), nlev(8);
  for (int o = 0; o < 8; o++) {
    const OctantPlan& op = p.oct[o];
    rank.insert(rank.end(), op.rank.begin(), op.rank.end());
    level.insert(level.end(), op.level.begin(), op.level.end());
    lvl_base[o] = (int)lo.size();
    pts_base[o] = (int)lp.size();
    nlev[o] = op.nlevels;
    lo.insert(lo.end(), op.lvl_off.begin(), op.lvl_off.end());
    lp.insert(lp.end(), op.lvl_pts.begin(), op.lvl_pts.end());
    lc.insert(lc.end(), op.lvl_cell.begin(), op.lvl_cell.end());
  }
  TRY(upload(s, s.rank, rank.data(), rank.size(), err));
  TRY(upload(s, s.level, level.data(), level.size(), err));
  TRY(upload(s, s.lvl_pts, lp.data(), lp.size(), err));
  TRY(upload(s, s.lvl_cell, lc.data(), lc.size(), err));
  TRY(upload(s, s.lvl_off, lo.data(), lo.size(), err));
  TRY(upload(s, s.oct_lvl_base, lvl_base.data(), 8, err));
  TRY(upload(s, s.oct_pts_base, pts_base.data(), 8, err));
  TRY(upload(s, s.oct_nlev, nlev.data(), 8, err));
  std::vector<float> r0(nsh, 0.f);
  for (int i = 0; i < p.npts; i++)
    for (int j = 0; j < 4 && j < s.ld; j++) r0[(size_t)i * s.ld + j] = p.radiance0[(size_t)i * 4 + j];
  TRY(upload(s, s.R0, r0.data(), r0.size(), err));
  for (int e = 0; e < State::NEV; e++) HIPCHK(hipEventCreate(&s.ev[e]));
  return AT3D_OK;
}

void state_free(State& s) {
  for (void* ptr : s.allocs) (void)hipFree(ptr);
  s.allocs.clear();
  for (int e = 0; e < State::NEV; e++)
    if (s.ev[e]) (void)hipEventDestroy(s.ev[e]);
  if (s.stream) (void)hipStreamDestroy(s.stream);
  s.stream = nullptr;
}

static SourceArgs source_args(State& s, int first, int fixsh) {
  const Plan& p = *s.plan;
  SourceArgs a;
  a.R = s.R;
  a.nr = s.nr;
  a.alb = s.alb;
  a.chi = s.chi;
  a.f0 = s.f0;
  a.ysun = s.ysun;
  a.lofj = s.lofj;
  a.S = s.S;
  a.DS = s.DS;
  a.ns = s.ns;
  a.nos = s.nos;
  a.partial = s.partial;
  a.ld = s.ld;
  a.npts = p.npts;
  a.ml = p.ml;
  a.mm = p.mm;
  a.nlm = p.nlm;
  a.shacc = p.d.shacc;
  a.first = first;
  a.accel = p.d.accelflag ? 1 : 0;
  a.fixsh = fixsh;
  return a;
}

int state_init_solution(State& s, std::string& err) {
  const Plan& p = *s.plan;
  const size_t nsh = (size_t)s.nptsp * s.ld;
  HIPCHK(hipMemcpyAsync(s.R, s.R0, nsh * sizeof(float), hipMemcpyDeviceToDevice, s.stream));
  std::vector<int32_t> four(p.npts, 4), zero(p.npts, 0);
  HIPCHK(hipMemcpyAsync(s.nr, four.data(), p.npts * sizeof(int), hipMemcpyHostToDevice, s.stream));
  HIPCHK(hipMemcpyAsync(s.ns, zero.data(), p.npts * sizeof(int), hipMemcpyHostToDevice, s.stream));
  HIPCHK(hipMemcpyAsync(s.nos, zero.data(), p.npts * sizeof(int), hipMemcpyHostToDevice, s.stream));
  SourceArgs a = source_args(s, 1, 0);
  k_compute_source<<<s.nsrc_blocks, 256, 0, s.stream>>>(a);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s.stream));
  s.solcrit = 1.0f;
  s.iters = 0;
  s.initialized = true;
  return AT3D_OK;
}

static SweepArgs sweep_args(State& s) {
  const Plan& p = *s.plan;
  SweepArgs a;
  a.g = GridTopo{s.gridpos, s.gridptr, s.neighptr, s.cellflags, p.npts, p.ncells};
  a.ordg = s.ordg;
  a.ord_oct = s.ordoct;
  a.skyrad = s.skyrad;
  a.bcrad_bot = s.bcrad_bot;
  a.bctop = s.bctop;
  a.bcbot = s.bcbot;
  a.ntop = (int)p.bctop.size();
  a.nbot = (int)p.bcbot.size();
  a.preset = s.preset;
  a.rank = s.rank;
  a.level = s.level;
  a.lvl_pts = s.lvl_pts;
  a.lvl_cell = s.lvl_cell;
  a.lvl_off = s.lvl_off;
  a.oct_lvl_base = s.oct_lvl_base;
  a.oct_pts_base = s.oct_pts_base;
  a.oct_nlev = s.oct_nlev;
  a.ext = s.ext;
  a.srcext = s.dosrc;
  a.rad = s.dorad;
  a.ldo = s.nptsp;
  a.npts = p.npts;
  a.errflag = s.errflag;
  return a;
}

// Host CALC_ACCEL_SOLCRIT (shdom_nompi.f:317-349) in single precision.
static void accel_solcrit(bool doaccel, float deljdot, float deljold, float deljnew, float jnorm,
                          float& a, float& accelpar, float& solcrit) {
  if (doaccel && a == 0.0f && deljnew < deljold) {
    float r = sqrtf(deljnew / deljold);
    float c = deljdot / sqrtf(deljold * deljnew);
    c = std::max(-1.0f, std::min(1.0f, c));
    float theta = acosf(c);
    a = (1 - r * cosf(theta) + powf(r, 1 + 0.5f * 3.14159f / theta)) /
            (1 + r * r - 2 * r * cosf(theta)) - 1.0f;
    a = std::min(10.0f, std::max(0.0f, a));
  } else {
    a = 0.0f;
  }
  accelpar = a;
  if (jnorm > 0.0f)
    solcrit = sqrtf(deljnew / jnorm);
  else if (deljnew == 0.0f)
    solcrit = 0.0f;
}

int state_solve(State& s, int maxiter, int verbose, std::string& err) {
  const Plan& p = *s.plan;
  if (!s.initialized) TRY(state_init_solution(s, err));
  HIPCHK(hipMemsetAsync(s.errflag, 0, sizeof(int), s.stream));
  float a = 0.0f;
  bool fixsh = false;
  const float radmin = p.d.shacc;
  SweepArgs sw = sweep_args(s);
  const dim3 g_s2d(s.nptsp / TN, s.nangp / TM);
  const int kpad = (int)round_up(p.nang, TK);
  const dim3 g_d2s(s.nptsp / TM, s.ld / TN);
  const int tb = 256, gpts = (p.npts + tb - 1) / tb;
  for (auto& t : s.tsum) t = 0.0;
  s.iters_last = 0;
  const float albpi = p.d.gndalbedo / acosf(-1.0f);
  while (s.iters < maxiter && s.solcrit > p.d.solacc) {
    s.iters++;
    s.iters_last++;
    (void)hipEventRecord(s.ev[0], s.stream);
    k_radiance_truncation<<<gpts, tb, 0, s.stream>>>(s.R, s.ld, s.tc, s.ns, s.lofj, s.nr, p.npts,
                                                     p.ml, p.mm, p.nlm, radmin, fixsh ? 1 : 0,
                                                     p.d.high_order_radiance ? 1 : 0);
    (void)hipEventRecord(s.ev[1], s.stream);
    k_sh_to_do<<<g_s2d, 256, 0, s.stream>>>(s.Y, s.S, s.ext, s.dosrc, s.ld, s.nptsp, p.nang, p.npts);
    (void)hipEventRecord(s.ev[2], s.stream);
    k_sweep<<<s.ndown, 1024, 0, s.stream>>>(sw, 0);
    k_bottom_bc<<<((int)p.bcbot.size() + 255) / 256, 256, 0, s.stream>>>(
        s.dorad, s.nptsp, s.bcbot, (int)p.bcbot.size(), s.fluxw, s.ndown, s.dirflux, albpi, s.bcrad_bot);
    k_sweep<<<p.nang - s.ndown, 1024, 0, s.stream>>>(sw, s.ndown);
    (void)hipEventRecord(s.ev[3], s.stream);
    k_do_to_sh<<<g_d2s, 256, 0, s.stream>>>(s.dorad, s.W, s.nr, s.R, kpad, s.nptsp, s.ld, p.npts);
    (void)hipEventRecord(s.ev[4], s.stream);
    if (s.solcrit < 0.001f || s.iters > p.d.iterfixsh) fixsh = true;
    SourceArgs sa = source_args(s, 0, fixsh ? 1 : 0);
    k_compute_source<<<s.nsrc_blocks, 256, 0, s.stream>>>(sa);
    k_reduce_dots<<<1, 256, 0, s.stream>>>(s.partial, s.nsrc_blocks, s.dots);
    (void)hipEventRecord(s.ev[5], s.stream);
    HIPCHK(hipGetLastError());
    double dots[4];
    HIPCHK(hipMemcpyAsync(dots, s.dots, sizeof(dots), hipMemcpyDeviceToHost, s.stream));
    HIPCHK(hipStreamSynchronize(s.stream));
    int ef = 0;
    HIPCHK(hipMemcpy(&ef, s.errflag, sizeof(int), hipMemcpyDeviceToHost));
    if (ef) {
      char buf[160];
      snprintf(buf, sizeof buf, "BACK_INT_GRID: invalid characteristic (flags 0x%x) at iteration %d",
               ef, s.iters);
      err = buf;
      return AT3D_ERR_NUMERIC;
    }
    float ms;
    for (int e = 0; e < 5; e++) {
      (void)hipEventElapsedTime(&ms, s.ev[e], s.ev[e + 1]);
      s.tsum[e] += ms;
    }
    s.deljdot = (float)dots[0];
    s.deljold = (float)dots[1];
    s.deljnew = (float)dots[2];
    s.jnorm = (float)dots[3];
    float accelpar;
    accel_solcrit(p.d.accelflag != 0, s.deljdot, s.deljold, s.deljnew, s.jnorm, a, accelpar, s.solcrit);
    if (accelpar > 0.0f) {
      const size_t n = (size_t)p.npts * s.ld;
      k_accelerate<<<(unsigned)((n + 255) / 256), 256, 0, s.stream>>>(s.S, s.DS, s.ns, s.nos, s.ld,
                                                                     p.npts, accelpar);
    }
    if (p.albmax < p.d.solacc) s.solcrit = p.d.solacc;
    if (verbose) {
      printf("  %4d %8.3f %8d\n", s.iters, log10f(std::max(s.solcrit, 1.0e-20f)), p.npts);
      fflush(stdout);
    }
  }
  k_fluxes<<<gpts, tb, 0, s.stream>>>(s.dorad, s.nptsp, s.fluxw, s.ndown, p.nang, p.npts, s.fluxes);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s.stream));
  return AT3D_OK;
}

}  // namespace at3d
