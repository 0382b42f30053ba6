// Spherical harmonic <-> discrete ordinate transforms as FP32 MFMA GEMMs.
//
// SH_TO_DO_UNPOL (shdomsub1.f:2928-3036) and DO_TO_SH_UNPOL (:3164-3254)
// factor the transform into a Legendre sum per m and an azimuthal FFT/DFT per
// zenith angle.  Here both are one dense matrix per direction (built on the
// host with the reference's m cut-off at NPHI0/2-1 per zenith angle), so each
// transform is a single [points x nlm] x [nlm x ordinates] product on the
// matrix cores.  The >=0 clamp of the DO source (PATH_INTEGRATION :2033-2037)
// and its product with the extinction are fused into the SH->DO epilogue; the
// per-point radiance truncation mask is fused into the DO->SH epilogue.
//
// Tiling: 64x64 output tile per 256-thread workgroup, K step 16, four waves in
// a 2x2 layout, each wave 2x2 blocks of v_mfma_f32_16x16x4_f32, whose operand
// lane layout is A[i][k] in lane i+16k, B[k][j] in lane j+16k, and
// D[4*(lane/16)+v][lane%16] in accumulator element v.  The LDS tiles are
// double-buffered: step k's tile is stored, one barrier, then step k+1's
// global loads are issued before step k's MFMAs, so they overlap.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace at3d {

typedef float float4_ __attribute__((ext_vector_type(4)));
constexpr int TM = 64, TN = 64, TK = 16, PAD = 4;

__device__ inline void mfma_tile(const float (*sA)[TM + PAD], const float (*sB)[TN + PAD], int wm,
                                 int wn, int lane, float4_ acc[2][2]) {
  // all of the step's LDS operands first, then the 16 MFMAs back to back
  float av[TK / 4][2], bv[TK / 4][2];
#pragma unroll
  for (int q = 0; q < TK / 4; q++) {
    const int kr = 4 * q + (lane >> 4);
#pragma unroll
    for (int i = 0; i < 2; i++) av[q][i] = sA[kr][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
    for (int j = 0; j < 2; j++) bv[q][j] = sB[kr][wn * 32 + j * 16 + (lane & 15)];
  }
#pragma unroll
  for (int q = 0; q < TK / 4; q++)
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q][i], bv[q][j], acc[i][j], 0, 0, 0);
}

// C[m][n] = sum_k A[m][k] B[n][k]: both operands K-contiguous.  CLAMP_EXT
// selects the SH->DO epilogue max(0, C) * ext[m] stored ordinate-major,
// out[n * ldo + m] (ldo = padded point count), four consecutive points per
// lane as one 16-byte store; without it C is stored as is, out[m * ldo + n].
// QU: K runs over two planes of A (Q, then U `plane` floats on) and B rows
// are 2 ld long.
template <bool CLAMP_EXT, bool QU = false>
__global__ __launch_bounds__(256) void k_sh_to_do(const float* __restrict__ S,
                                                  const float* __restrict__ Yc,
                                                  const float* __restrict__ ext,
                                                  float* __restrict__ out, int ld, int ldo,
                                                  int npts, int nang, const int* __restrict__ stop,
                                                  size_t plane = 0) {
  if (stop && *stop) return;
  __shared__ float sA[2][TK][TM + PAD];
  __shared__ float sB[2][TK][TN + PAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  float4_ acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = (float4_){0.f, 0.f, 0.f, 0.f};
  const int lr = tid >> 2, lc = (tid & 3) * 4;
  const int kt = QU ? 2 * ld : ld;   // K extent and B row length
  const float* arow = S + (size_t)(m0 + lr) * ld + lc;
  const float* brow = Yc + (size_t)(n0 + lr) * kt + lc;
  // K index k of A: the Q plane below ld, the U plane above
  auto aptr = [&](int k) { return (QU && k >= ld) ? arow + plane + (k - ld) : arow + k; };
  // double-buffered LDS: the next K step's global loads fly during this step's MFMAs
  float4 a = *reinterpret_cast<const float4*>(arow);
  float4 b = *reinterpret_cast<const float4*>(brow);
  for (int k0 = 0, buf = 0; k0 < kt; k0 += TK, buf ^= 1) {
    sA[buf][lc + 0][lr] = a.x;
    sA[buf][lc + 1][lr] = a.y;
    sA[buf][lc + 2][lr] = a.z;
    sA[buf][lc + 3][lr] = a.w;
    sB[buf][lc + 0][lr] = b.x;
    sB[buf][lc + 1][lr] = b.y;
    sB[buf][lc + 2][lr] = b.z;
    sB[buf][lc + 3][lr] = b.w;
    __syncthreads();
    if (k0 + TK < kt) {
      a = *reinterpret_cast<const float4*>(aptr(k0 + TK));
      b = *reinterpret_cast<const float4*>(brow + k0 + TK);
    }
    mfma_tile(sA[buf], sB[buf], wm, wn, lane, acc);
  }
  if (CLAMP_EXT) {
    // rows beyond npts hold zero sources and zero extinction: stored as zeros
    for (int i = 0; i < 2; i++) {
      const int p0 = m0 + wm * 32 + i * 16 + 4 * (lane >> 4);
      const float4 e = *reinterpret_cast<const float4*>(ext + p0);
      for (int j = 0; j < 2; j++) {
        const int c = n0 + wn * 32 + j * 16 + (lane & 15);
        if (c >= nang) continue;
        float4 o;
        if (QU) {   // Q and U are not clamped
          o.x = acc[i][j][0] * e.x;
          o.y = acc[i][j][1] * e.y;
          o.z = acc[i][j][2] * e.z;
          o.w = acc[i][j][3] * e.w;
        } else {
          o.x = fmaxf(acc[i][j][0], 0.0f) * e.x;
          o.y = fmaxf(acc[i][j][1], 0.0f) * e.y;
          o.z = fmaxf(acc[i][j][2], 0.0f) * e.z;
          o.w = fmaxf(acc[i][j][3], 0.0f) * e.w;
        }
        *reinterpret_cast<float4*>(out + (size_t)c * ldo + p0) = o;
      }
    }
  } else {
    for (int i = 0; i < 2; i++)
      for (int v = 0; v < 4; v++) {
        const int p = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + v;
        if (p >= npts) continue;
        for (int j = 0; j < 2; j++) {
          const int c = n0 + wn * 32 + j * 16 + (lane & 15);
          if (c < nang) out[(size_t)p * ldo + c] = acc[i][j][v];
        }
      }
  }
}

// SH->DO in the sweep's level order: one workgroup computes 64 consecutive
// positions of one octant's order (S rows gathered through perm) for up to 48
// columns of that octant; each wave owns 16 positions x 48 columns (three
// 16x16 MFMA blocks).  S is zero beyond each point's source truncation ns
// (COMPUTE_SOURCE keeps NS terms, shdomsub1.f:967-1611), so the K loop stops
// at the largest ns of the tile's rows, rounded up to TK: the terms skipped
// are exact zeros, and the sweep's order groups points of like ns.
// QU: the polarized Q / U rows, K over the Q plane's kn terms then the U
// plane's (`plane` floats on), B rows 2 ld long, no clamp at zero.
constexpr int PN = 48;
template <bool QU>
__global__ __launch_bounds__(256) void k_sh_to_do_perm(const float* __restrict__ S,
                                                       const float* __restrict__ Yc,
                                                       const float* __restrict__ ext,
                                                       const int* __restrict__ ns,
                                                       const int32_t* __restrict__ oct_perm,
                                                       const int4* __restrict__ tiles,
                                                       float* __restrict__ out, int ld, int ldp,
                                                       int npts, const int* __restrict__ stop,
                                                       size_t plane) {
  if (stop && *stop) return;
  __shared__ float sA[2][TK][TM + PAD];
  __shared__ float sB[2][TK][PN + PAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int4 tile = tiles[blockIdx.y];
  const int32_t* perm = oct_perm + (size_t)tile.x * ldp;
  const int c0 = tile.y, nc = tile.z;
  const int m0 = blockIdx.x * TM;
  // every wave reduces the same 64 rows: the K extent is uniform in the block
  int kn;
  {
    const int pr = perm[m0 + lane];
    int v = pr < npts ? ns[pr] : 0;
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    kn = __builtin_amdgcn_readfirstlane(min(ld, (v + TK - 1) / TK * TK));
  }
  float4_ acc[3];
  for (int j = 0; j < 3; j++) acc[j] = (float4_){0.f, 0.f, 0.f, 0.f};
  const int lr = tid >> 2, lc = (tid & 3) * 4;
  const float* arow = S + (size_t)perm[m0 + lr] * ld + lc;
  const bool bload = lr < nc;  // threads 0..4*nc-1 load the column tile
  const float* brow = Yc + (size_t)(c0 + (bload ? lr : 0)) * (QU ? 2 * ld : ld) + lc;
  // step kk of the K loop: term kk of the Q plane, or kk - kn of the U plane
  const int kend = QU ? 2 * kn : kn;
  auto aptr = [&](int kk) { return (QU && kk >= kn) ? arow + plane + (kk - kn) : arow + kk; };
  auto bptr = [&](int kk) { return (QU && kk >= kn) ? brow + ld + (kk - kn) : brow + kk; };
  float4 a = *reinterpret_cast<const float4*>(arow);
  float4 b = make_float4(0.f, 0.f, 0.f, 0.f);
  if (bload) b = *reinterpret_cast<const float4*>(brow);
  for (int k0 = 0, buf = 0; k0 < kend; k0 += TK, buf ^= 1) {
    sA[buf][lc + 0][lr] = a.x;
    sA[buf][lc + 1][lr] = a.y;
    sA[buf][lc + 2][lr] = a.z;
    sA[buf][lc + 3][lr] = a.w;
    if (lr < PN) {
      sB[buf][lc + 0][lr] = b.x;
      sB[buf][lc + 1][lr] = b.y;
      sB[buf][lc + 2][lr] = b.z;
      sB[buf][lc + 3][lr] = b.w;
    }
    __syncthreads();
    if (k0 + TK < kend) {
      a = *reinterpret_cast<const float4*>(aptr(k0 + TK));
      if (bload) b = *reinterpret_cast<const float4*>(bptr(k0 + TK));
    }
    float av[TK / 4], bv[TK / 4][3];
#pragma unroll
    for (int q = 0; q < TK / 4; q++) {
      const int kr = 4 * q + (lane >> 4);
      av[q] = sA[buf][kr][wave * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 3; j++) bv[q][j] = sB[buf][kr][j * 16 + (lane & 15)];
    }
#pragma unroll
    for (int q = 0; q < TK / 4; q++)
#pragma unroll
      for (int j = 0; j < 3; j++)
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q], bv[q][j], acc[j], 0, 0, 0);
  }
  // rows beyond npts gather the zero padding rows of S and ext
  const int q0 = m0 + wave * 16 + 4 * (lane >> 4);
  float e[4];
  for (int v = 0; v < 4; v++) e[v] = ext[perm[q0 + v]];
  for (int j = 0; j < 3; j++) {
    const int c = j * 16 + (lane & 15);
    if (c >= nc) continue;
    float4 o;
    if (QU) {
      o.x = acc[j][0] * e[0];
      o.y = acc[j][1] * e[1];
      o.z = acc[j][2] * e[2];
      o.w = acc[j][3] * e[3];
    } else {
      o.x = fmaxf(acc[j][0], 0.0f) * e[0];
      o.y = fmaxf(acc[j][1], 0.0f) * e[1];
      o.z = fmaxf(acc[j][2], 0.0f) * e[2];
      o.w = fmaxf(acc[j][3], 0.0f) * e[3];
    }
    *reinterpret_cast<float4*>(out + (size_t)(c0 + c) * ldp + q0) = o;
  }
}

// C[m][n] = sum_k A[m][k] B[k][n]: A stored ordinate-major (A[m][k] at
// D[k * ldp + m]), B N-contiguous rows.  QU: K over 2 ldo rows (the Q rows,
// then the U rows that follow them), B rows 2 ld long, output column n into
// plane n / ld of R (R_Q, then R_U `plane` floats on) at term n % ld.
template <bool QU>
__global__ __launch_bounds__(256) void k_do_to_sh(const float* __restrict__ D,
                                                  const float* __restrict__ Wc,
                                                  const int* __restrict__ nr,
                                                  float* __restrict__ R, int ld, int ldo, int ldp,
                                                  int npts, const int* __restrict__ stop, size_t plane) {
  if (stop && *stop) return;
  __shared__ float sA[2][TK][TM + PAD];
  __shared__ float sB[2][TK][TN + PAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  // XCD-aware tile order: workgroups are dealt to the eight XCDs round robin
  // (each with its own L2), so the ld/TN SH tiles of one point tile are given
  // workgroup ids that land on one XCD in a row; the point tile's DO
  // radiances are then read from HBM once and from that XCD's L2 after.
  const int ntn = gridDim.x, nwg = gridDim.x * gridDim.y;
  const int bid = blockIdx.y * gridDim.x + blockIdx.x;
  int mt, nt;
  if (nwg % (8 * ntn) == 0) {
    const int xcd = bid & 7, local = bid >> 3;
    mt = (local / ntn) * 8 + xcd;
    nt = local % ntn;
  } else {
    mt = blockIdx.y;
    nt = blockIdx.x;
  }
  const int m0 = mt * TM;
  const int ldb = QU ? 2 * ld : ld, kext = QU ? 2 * ldo : ldo;
  const int nb = nt * TN;                  // column of B
  const int pl = QU ? nb / ld : 0;         // output plane
  const int n0 = nb - pl * ld;             // first SH term of the tile
  R += pl * plane;
  // an SH tile at or past every radiance count of the point tile is zero:
  // written as such without the product (radiance truncation leaves most
  // clear-sky points a few low-order terms)
  __shared__ int tile_nr;
  if (wave == 0) {
    int v = (m0 + lane < npts) ? nr[m0 + lane] : 0;
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    if (lane == 0) tile_nr = v;
  }
  __syncthreads();
  if (n0 >= tile_nr) {
    for (int i = tid; i < TM * (TN / 4); i += 256) {
      const int r = i / (TN / 4), c = (i - r * (TN / 4)) * 4;
      if (m0 + r < npts)
        *reinterpret_cast<float4*>(R + (size_t)(m0 + r) * ld + n0 + c) = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    return;
  }
  float4_ acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = (float4_){0.f, 0.f, 0.f, 0.f};
  const int bk = tid >> 4, bc = (tid & 15) * 4;    // 16 k x 64 rows/cols per tile
  const float* acol = D + (size_t)bk * ldp + m0 + bc;
  const float* bcol = Wc + (size_t)bk * ldb + nb + bc;
  float4 a = *reinterpret_cast<const float4*>(acol);
  float4 b = *reinterpret_cast<const float4*>(bcol);
  for (int k0 = 0, buf = 0; k0 < kext; k0 += TK, buf ^= 1) {
    *reinterpret_cast<float4*>(&sA[buf][bk][bc]) = a;
    *reinterpret_cast<float4*>(&sB[buf][bk][bc]) = b;
    __syncthreads();
    if (k0 + TK < kext) {
      a = *reinterpret_cast<const float4*>(acol + (size_t)(k0 + TK) * ldp);
      b = *reinterpret_cast<const float4*>(bcol + (size_t)(k0 + TK) * ldb);
    }
    mfma_tile(sA[buf], sB[buf], wm, wn, lane, acc);
  }
  // the eight rows' radiance counts first (independent loads), then the stores
  int n[2][4];
  for (int i = 0; i < 2; i++)
    for (int v = 0; v < 4; v++) {
      const int p = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + v;
      n[i][v] = p < npts ? nr[p] : -1;
    }
  for (int i = 0; i < 2; i++)
    for (int v = 0; v < 4; v++) {
      if (n[i][v] < 0) continue;
      const int p = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + v;
      for (int j = 0; j < 2; j++) {
        const int J = n0 + wn * 32 + j * 16 + (lane & 15);
        R[(size_t)p * ld + J] = (J < n[i][v]) ? acc[i][j][v] : 0.0f;
      }
    }
}

void launch_sh_to_do(hipStream_t st, const float* S, const float* Yc, const float* ext,
                     float* dosrc, int mp, int ncp, int ld, int ldp, int npts, int nang,
                     const int* stop) {
  dim3 grid(ncp / TN, mp / TM);
  k_sh_to_do<true><<<grid, 256, 0, st>>>(S, Yc, ext, dosrc, ld, ldp, npts, nang, stop);
}

void launch_sh_to_do_qu(hipStream_t st, const float* S, size_t plane, const float* Yc, const float* ext,
                        float* dosrc, int mp, int ncp, int ld, int ldp, int npts, const int* stop) {
  dim3 grid(ncp / TN, mp / TM);
  k_sh_to_do<true, true><<<grid, 256, 0, st>>>(S, Yc, ext, dosrc, ld, ldp, npts, ncp, stop, plane);
}

void launch_sh_dot(hipStream_t st, const float* S, const float* Yd, float* raw, int mp, int ndp,
                   int ld, int ldd, int npts, int ndir) {
  dim3 grid(ndp / TN, mp / TM);
  k_sh_to_do<false><<<grid, 256, 0, st>>>(S, Yd, nullptr, raw, ld, ldd, npts, ndir, nullptr);
}

void launch_sh_to_do_perm(hipStream_t st, const float* S, const float* Yc, const float* ext,
                          const int* ns, const int32_t* oct_perm, const int4* tiles, int ntiles,
                          float* dosrc, int mp, int ld, int ldp, int npts, const int* stop) {
  if (ntiles <= 0) return;
  dim3 grid(mp / TM, ntiles);
  k_sh_to_do_perm<false><<<grid, 256, 0, st>>>(S, Yc, ext, ns, oct_perm, tiles, dosrc, ld, ldp, npts, stop, 0);
}

void launch_sh_to_do_perm_qu(hipStream_t st, const float* S, size_t plane, const float* Yc, const float* ext,
                             const int* ns, const int32_t* oct_perm, const int4* tiles, int ntiles,
                             float* dosrc, int mp, int ld, int ldp, int npts, const int* stop) {
  if (ntiles <= 0) return;
  dim3 grid(mp / TM, ntiles);
  k_sh_to_do_perm<true><<<grid, 256, 0, st>>>(S, Yc, ext, ns, oct_perm, tiles, dosrc, ld, ldp, npts, stop,
                                              plane);
}

void launch_do_to_sh(hipStream_t st, const float* dorad, const float* Wc, const int* nr, float* R,
                     int mp, int ld, int ldo, int ldp, int npts, const int* stop) {
  dim3 grid(ld / TN, mp / TM);
  k_do_to_sh<false><<<grid, 256, 0, st>>>(dorad, Wc, nr, R, ld, ldo, ldp, npts, stop, 0);
}

void launch_do_to_sh_qu(hipStream_t st, const float* dorad, const float* Wc, const int* nr, float* R, size_t plane,
                        int mp, int ld, int ldo, int ldp, int npts, const int* stop) {
  dim3 grid(2 * ld / TN, mp / TM);
  k_do_to_sh<true><<<grid, 256, 0, st>>>(dorad, Wc, nr, R, ld, ldo, ldp, npts, stop, plane);
}

void launch_sh_dot_qu(hipStream_t st, const float* S, size_t plane, const float* Yd, float* raw, int mp,
                      int ndp, int ld, int ldo, int npts) {
  dim3 grid(ndp / TN, mp / TM);
  k_sh_to_do<false, true><<<grid, 256, 0, st>>>(S, Yd, nullptr, raw, ld, ldo, npts, ndp, nullptr, plane);
}

}  // namespace at3d
