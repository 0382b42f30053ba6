// Source function, spherical-harmonic truncation and sequence acceleration.
#include <hip/hip_runtime.h>
#include <math.h>

#include "kernels.h"

namespace at3d {

__device__ inline int nterms_for_l(int l, int mm) {
  return (l <= mm) ? l * (l + 1) + l + 1 : (2 * mm + 1) * l - mm * mm + mm + 1;
}

// RADIANCE_TRUNCATION (shdomsub1.f:1615-1803): per-point radiance SH count.
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

// COMPUTE_SOURCE for NSTOKES=1 (shdomsub1.f:967-1611).  One wavefront per grid
// point (SRC_PTS_PER_BLOCK / 4 points per wave in turn), up to eight terms per
// lane (NLM <= 512): every load of a point is issued before its arithmetic.
// SOURCET_J = w chi_l (F0 Ysun_J + I_J); the acceleration dot products
// accumulate in FP64 per lane in term order and are written as one partial per
// block for a fixed-order reduction.
// NT terms per lane: NLM <= 64 * NT (4 up to L = 15, 8 up to NLM 512; the
// smaller instantiation holds fewer registers and keeps twice the waves in
// flight).  NST = 3: the polarized source (CALC_SOURCE_PNT, shdomsub1.f:
// 823-910) of the I, Q and U planes from the a1, a2, a3 and b1 terms,
//   S^I_J = w (a1_l R^I_J + b1_l R^Q_J) + F0 w a1_l Ysun_J
//   S^Q_J = w (b1_l R^I_J + a2_l R^Q_J) + F0 w b1_l Ysun_J    (l >= 2)
//   S^U_J = w a3_l R^U_J                                       (l >= 2)
// the truncation over all three, the dot products summed over them.
template <int NT, int NST>
__global__ __launch_bounds__(256) void k_compute_source(SourceArgs a) {
  if (a.ctl && a.ctl->stop) return;
  const int fixsh = a.ctl ? a.ctl->fixsh : a.fixsh;
  constexpr int SRC_TERMS = NT;
  __shared__ double red[4][4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double dot = 0, dold = 0, dnew = 0, jn = 0;
  // l(J) and Ysun_J of this lane's terms, the same for every point
  int lv[SRC_TERMS];
  float yv[SRC_TERMS];
#pragma unroll
  for (int it = 0; it < SRC_TERMS; it++) {
    const int J = lane + 64 * it;
    lv[it] = J < a.nlm ? a.lofj[J] : 0;
    yv[it] = J < a.nlm ? a.ysun[J] : 0.f;
  }
  // a point's scalars (albedo, F0, NR, NS, NOS) are loaded one point ahead,
  // so the row loads of a point wait for one memory round trip, not two
  const int p0 = blockIdx.x * SRC_PTS_PER_BLOCK + w;
  float alb_n = 0.f, f0_n = 0.f;
  int nr_n = 0, ns_n = 0, nos_n = 0;
  if (p0 < a.npts) {
    alb_n = a.alb[p0];
    f0_n = a.f0[p0];
    nr_n = a.nr[p0];
    ns_n = a.ns[p0];
    nos_n = a.nos[p0];
  }
  for (int sub = 0; sub < SRC_PTS_PER_BLOCK / 4; sub++) {
    const int p = p0 + sub * 4;
    if (p >= a.npts) break;
    const float alb = alb_n, f0 = f0_n;
    const int nrp = nr_n, ns_old = ns_n, nos_old = nos_n;
    if (sub + 1 < SRC_PTS_PER_BLOCK / 4 && p + 4 < a.npts) {
      alb_n = a.alb[p + 4];
      f0_n = a.f0[p + 4];
      nr_n = a.nr[p + 4];
      ns_n = a.ns[p + 4];
      nos_n = a.nos[p + 4];
    }
    const float* Rp = a.R + (size_t)p * a.ld;
    float* Sp = a.S + (size_t)p * a.ld;
    float* DSp = a.DS + (size_t)p * a.ld;
    const float* chip = a.chi + (size_t)p * (a.ml + 1);
    const float* chpp = NST == 3 ? a.chip + (size_t)p * 3 * (a.ml + 1) : nullptr;
    const bool old = !a.first;
    const bool useds = old && a.accel;
    const int nsd = a.accel ? min(ns_old, nos_old) : ns_old;
    // the truncation constants of the next radiance truncation, loaded with
    // the point's rows (the first 64 orders)
    const float* tcp = a.tc + (size_t)p * (a.ml + 1);
    const float tcv = (!fixsh && 1 + lane <= a.ml) ? tcp[1 + lane] : 0.f;
    // per Stokes plane k: radiance, old source, delta source; the phase terms
    // a1 (and a2, a3, b1)
    float rv[NST][SRC_TERMS], sv[NST][SRC_TERMS], dv[NST][SRC_TERMS], cv[SRC_TERMS];
    float cp[NST == 3 ? 3 : 1][SRC_TERMS];
#pragma unroll
    for (int it = 0; it < SRC_TERMS; it++) {
      const int J = lane + 64 * it;
      cv[it] = 0.f;
#pragma unroll
      for (int k = 0; k < NST; k++) rv[k][it] = sv[k][it] = dv[k][it] = 0.f;
#pragma unroll
      for (int e = 0; e < (NST == 3 ? 3 : 1); e++) cp[e][it] = 0.f;
      if (J < a.nlm) {
        cv[it] = chip[lv[it]];
        if constexpr (NST == 3) {
#pragma unroll
          for (int e = 0; e < 3; e++) cp[e][it] = chpp[e * (a.ml + 1) + lv[it]];
        }
#pragma unroll
        for (int k = 0; k < NST; k++)
          if (J < nrp) rv[k][it] = Rp[k * a.psh + J];
      }
      // S is zero beyond NS and DS beyond NOS (both kept so below): only
      // the terms that can be nonzero are read
#pragma unroll
      for (int k = 0; k < NST; k++) {
        if (old && J < ns_old) sv[k][it] = Sp[k * a.psh + J];
        if (useds && J < nsd) dv[k][it] = DSp[k * a.psh + J];
      }
    }
    int js = -1;
    float st[NST][SRC_TERMS];
#pragma unroll
    for (int it = 0; it < SRC_TERMS; it++) {
      const int J = lane + 64 * it;
      if constexpr (NST == 1) {
        float v = 0.f;
        if (J < a.nlm) {
          const float c = cv[it];
          v = f0 * alb * c * yv[it];
          if (J < nrp) v = v + alb * c * rv[0][it];
          if (fabsf(v) > a.shacc) js = J;
        }
        st[0][it] = v;
      } else {
        float vi = 0.f, vq = 0.f, vu = 0.f;
        if (J < a.nlm) {
          const float c1 = cv[it], c2 = cp[0][it], c3 = cp[1][it], c5 = cp[2][it];
          if (J < nrp) vi = alb * (c1 * rv[0][it] + c5 * rv[1][it]);
          vi = vi + f0 * alb * c1 * yv[it];
          if (J >= 4) {
            if (J < nrp) {
              vq = alb * (c5 * rv[0][it] + c2 * rv[1][it]);
              vu = alb * (c3 * rv[2][it]);
            }
            vq = vq + f0 * alb * c5 * yv[it];
          }
          if (fabsf(vi) > a.shacc || fabsf(vq) > a.shacc || fabsf(vu) > a.shacc) js = J;
        }
        st[0][it] = vi;
        st[1][it] = vq;
        st[2][it] = vu;
      }
    }
    if (old) {
#pragma unroll
      for (int it = 0; it < SRC_TERMS; it++) {
        const int J = lane + 64 * it;
        if (J < nsd) {
#pragma unroll
          for (int k = 0; k < NST; k++) {
            const double so = sv[k][it];
            const double dn = (double)(st[k][it] - sv[k][it]);
            if (a.accel) {
              const double d = dv[k][it];
              dot += dn * d;
              dold += d * d;
            }
            dnew += dn * dn;
            jn += so * so;
          }
        }
      }
      if (a.accel) {
        // DS was zero from the previous NOS on: terms up to the larger of the
        // two counts are written
        const int dw = max(ns_old, nos_old);
#pragma unroll
        for (int it = 0; it < SRC_TERMS; it++) {
          const int J = lane + 64 * it;
#pragma unroll
          for (int k = 0; k < NST; k++)
            if (J < dw) DSp[k * a.psh + J] = (J < ns_old) ? st[k][it] - sv[k][it] : 0.f;
        }
        if (lane == 0) a.nos[p] = ns_old;
      }
    }
    int nsn;
    if (fixsh) {
      nsn = ns_old;
    } else {
      for (int off = 32; off > 0; off >>= 1) js = max(js, __shfl_xor(js, off));
      nsn = (js < 0) ? 0 : nterms_for_l(a.lofj[js], a.mm);
    }
    // terms below the larger of the old and new NS (beyond both the row is
    // zero already; in the first pass NS is the extent of an earlier solve's
    // terms)
    const int sw = max(nsn, ns_old);
#pragma unroll
    for (int it = 0; it < SRC_TERMS; it++) {
      const int J = lane + 64 * it;
#pragma unroll
      for (int k = 0; k < NST; k++)
        if (J < sw) Sp[k * a.psh + J] = (J < nsn && J < a.nlm) ? st[k][it] : 0.f;
    }
    if (lane == 0) a.ns[p] = nsn;
    // the next iteration's radiance truncation (RADIANCE_TRUNCATION
    // shdomsub1.f:1615-1803, as k_radiance_truncation): it reads NS as just
    // written and the radiance's first term (lane 0's first term here)
    int nrn;
    if (fixsh) {
      nrn = a.highorder ? nterms_for_l(a.ml, a.mm) : max(4, nsn);
    } else {
      const float i0 = __shfl(rv[0][0], 0);
      int lr = 1;
      if (1 + lane <= a.ml && tcv * i0 > a.radmin) lr = 1 + lane;
      for (int l = 65 + lane; l <= a.ml; l += 64)   // ML >= 64 (M much below L)
        if (tcp[l] * i0 > a.radmin) lr = l;
      for (int off = 32; off > 0; off >>= 1) lr = max(lr, __shfl_xor(lr, off));
      const int ls = a.lofj[max(1, nsn) - 1];
      lr = min(lr, ls + a.ml / 8 + 2);
      if (a.highorder) lr = a.ml;
      nrn = nterms_for_l(lr, a.mm);
    }
    if (lane == 0) a.nr_next[p] = nrn;
    if (a.first && a.accel) {
      // DS cleared (it is zero from NOS on)
#pragma unroll
      for (int it = 0; it < SRC_TERMS; it++) {
        const int J = lane + 64 * it;
#pragma unroll
        for (int k = 0; k < NST; k++)
          if (J < nos_old) DSp[k * a.psh + J] = 0.f;
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

// INIT_RADIANCE's terms into the SH radiance rows: one wavefront per point.
__global__ __launch_bounds__(256) void k_init_radiance(float* __restrict__ R, const float* __restrict__ R0,
                                                       int* __restrict__ nr, int ld, int npts) {
  const int lane = threadIdx.x & 63;
  const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= npts) return;
  const int n = min(ld, max(4, nr[p]));
  float* Rp = R + (size_t)p * ld;
  for (int J = lane; J < n; J += 64) Rp[J] = J < 4 ? R0[(size_t)p * 4 + J] : 0.f;
  if (lane == 0) nr[p] = 4;
}

// Fixed-order reduction of the block partials.
__global__ __launch_bounds__(1024) void k_reduce_dots(const double* __restrict__ partial, int nblocks,
                                                      double* __restrict__ out,
                                                      const int* __restrict__ stop) {
  if (stop && *stop) return;
  __shared__ double sh[1024];
  const int c = threadIdx.x & 3, t = threadIdx.x >> 2;  // 256 threads per component
  double s = 0;
  for (int b = t; b < nblocks; b += 256) s += partial[(size_t)b * 4 + c];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int stride = 512; stride >= 4; stride >>= 1) {
    if ((int)threadIdx.x < stride) sh[threadIdx.x] += sh[threadIdx.x + stride];
    __syncthreads();
  }
  if (threadIdx.x < 4) out[threadIdx.x] = sh[threadIdx.x];
}

// ACCELERATE_SOLUTION (shdomsub1.f:1807-1831): one wavefront per point over
// its first min(NS, NOS) terms.
__global__ __launch_bounds__(256) void k_accelerate(float* __restrict__ S, const float* __restrict__ DS,
                                                    const int* __restrict__ ns, const int* __restrict__ nos,
                                                    int ld, int npts, float a,
                                                    const IterCtl* __restrict__ ctl, size_t psh) {
  if (ctl) {
    if (ctl->stop || !(ctl->a > 0.0f)) return;
    a = ctl->a;
  }
  const int lane = threadIdx.x & 63;
  const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= npts) return;
  const int n = min(ns[p], nos[p]);
  // blockIdx.y: the Stokes plane
  float* Sp = S + blockIdx.y * psh + (size_t)p * ld;
  const float* Dp = DS + blockIdx.y * psh + (size_t)p * ld;
  for (int J = lane; J < n; J += 64) Sp[J] = Sp[J] + a * Dp[J];
}

void launch_radiance_truncation(hipStream_t st, const float* R, int ld, const float* tc,
                                const int* ns, const int* lofj, int* nr, int npts, int ml, int mm,
                                int nlm, float radmin, int fixsh, int highorder) {
  k_radiance_truncation<<<(npts + 255) / 256, 256, 0, st>>>(R, ld, tc, ns, lofj, nr, npts, ml, mm,
                                                             nlm, radmin, fixsh, highorder);
}

void launch_compute_source(hipStream_t st, const SourceArgs& a, int nblocks) {
  if (a.nst == 3) {
    if (a.nlm <= 256)
      k_compute_source<4, 3><<<nblocks, 256, 0, st>>>(a);
    else
      k_compute_source<8, 3><<<nblocks, 256, 0, st>>>(a);
    return;
  }
  if (a.nlm <= 256)
    k_compute_source<4, 1><<<nblocks, 256, 0, st>>>(a);
  else
    k_compute_source<8, 1><<<nblocks, 256, 0, st>>>(a);
}

void launch_init_radiance(hipStream_t st, float* R, const float* R0, int* nr, int ld, int npts) {
  k_init_radiance<<<(npts + 3) / 4, 256, 0, st>>>(R, R0, nr, ld, npts);
}

void launch_reduce_dots(hipStream_t st, const double* partial, int nblocks, double* out,
                        const int* stop) {
  k_reduce_dots<<<1, 1024, 0, st>>>(partial, nblocks, out, stop);
}

void launch_accelerate(hipStream_t st, float* S, const float* DS, const int* ns, const int* nos,
                       int ld, int npts, float a, const IterCtl* ctl, int nst, size_t psh) {
  k_accelerate<<<dim3((npts + 3) / 4, nst), 256, 0, st>>>(S, DS, ns, nos, ld, npts, a, ctl, psh);
}

// The loop test of SOLUTION_ITERATIONS (shdomsub1.f:674) and FIXSH
// (:679-681; FIXSH once SOLCRIT < 0.001 or past ITERFIXSH iterations).
__global__ void k_iter_begin(IterCtl* ctl) {
  if (threadIdx.x != 0 || ctl->stop) return;
  if (ctl->total >= ctl->maxiter || ctl->solcrit <= ctl->solacc) {
    ctl->stop = 1;
    return;
  }
  ctl->total = ctl->total + 1;
  ctl->iters = ctl->iters + 1;
  if (ctl->solcrit < 0.001f || ctl->total > ctl->iterfixsh) ctl->fixsh = 1;
}

// CALC_ACCEL_SOLCRIT (shdom_nompi.f:317-349) in single precision from the
// reduced FP64 dot products: the acceleration parameter of this iteration
// (alternating: none after an accelerated iteration) and SOLCRIT; with
// ALBMAX < SOLACC one iteration is enough (shdomsub1.f:787).  An error flag
// raised by a kernel of the iteration stops the solve.
__global__ void k_solcrit(IterCtl* ctl, const double* __restrict__ dots, const int* __restrict__ errflag) {
  if (threadIdx.x != 0 || ctl->stop) return;
  const float deljdot = (float)dots[0], deljold = (float)dots[1], deljnew = (float)dots[2],
              jnorm = (float)dots[3];
  float a = ctl->a;
  if (ctl->accel && a == 0.0f && deljnew < deljold) {
    const float r = sqrtf(deljnew / deljold);
    float c = deljdot / sqrtf(deljold * deljnew);
    c = fmaxf(-1.0f, fminf(1.0f, c));
    const float theta = acosf(c);
    a = (1 - r * cosf(theta) + powf(r, 1 + 0.5f * 3.14159f / theta)) / (1 + r * r - 2 * r * cosf(theta)) -
        1.0f;
    a = fminf(10.0f, fmaxf(0.0f, a));
  } else {
    a = 0.0f;
  }
  ctl->a = a;
  float solcrit = ctl->solcrit;
  if (jnorm > 0.0f)
    solcrit = sqrtf(deljnew / jnorm);
  else if (deljnew == 0.0f)
    solcrit = 0.0f;
  if (ctl->one_iter) solcrit = ctl->solacc;
  ctl->solcrit = solcrit;
  ctl->deljdot = deljdot;
  ctl->deljold = deljold;
  ctl->deljnew = deljnew;
  ctl->jnorm = jnorm;
  if (*errflag) ctl->stop = 2;
}

void launch_iter_begin(hipStream_t st, IterCtl* ctl) { k_iter_begin<<<1, 64, 0, st>>>(ctl); }

void launch_solcrit(hipStream_t st, IterCtl* ctl, const double* dots, const int* errflag) {
  k_solcrit<<<1, 64, 0, st>>>(ctl, dots, errflag);
}

}  // namespace at3d
