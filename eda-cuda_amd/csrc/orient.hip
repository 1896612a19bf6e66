// SPDX-License-Identifier: Apache-2.0
// The monogenic orientation of a mode (reference cpp/src/method/orientation.cu:1-26,136-404,
// element.cu:29-164; long-double twin cpp/tests/twin_orientation.hpp:1-120):
//   X = the orthonormal N-D DCT-II of r - mean(r), X_b = G(rad) X with the one-octave log-Gabor
//   band at f0 = 1 / w (G(0) = 0), the Riesz components of the band by the inverse transform with
//   the sin basis on the odd axes:
//     R1_a  = sum_k c X_b (u_a / rad) sin(theta_a) prod_{b != a} cos(theta_b)
//     R2_aa = sum_k c X_b (-(u_a / rad)^2) prod_b cos(theta_b)
//     R2_ab = sum_k c X_b (u_a u_b / rad^2) sin(theta_a) sin(theta_b) prod_{others} cos(theta)
//   (u_a = k_a / (2 n_a), theta_a = pi k_a (2 x_a + 1) / (2 n_a), c = prod_a c_{k_a}),
//   T = R1 R1^T + R2 R2^T per voxel, and its Jacobi eigen-decomposition into the top eigenvector n
//   and the coherence c = (mu1 - mu2) / (mu1 + mu2).
// The reference computes the transforms with cuFFT through Makhoul's permutation. Here a
// power-of-two axis runs Makhoul's reordering and an LDS Stockham FFT (k_dct_fft); other axes are
// contracted with their dense DCT matrix as batched GEMMs (the DCT of an n-sample axis is an n x n
// matrix, so a transform of a rows x n block is one GEMM); dct="bf16x3" contracts in split-BF16 MFMA
// products. The matrices and twiddles are built on the device from cos / sin tables of
// pi j / (2n), j < 4n, evaluated in long double on the host.
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace eda {
namespace {

// ---- the batched GEMM -----------------------------------------------------------------------------
// C[bt](i, j) = sum_k A[bt](i, k) B[bt](k, j) with element strides; batch bt runs when
// on[bt / per_elem] != 0. The float64 kernel: 64 x 64 tiles, K steps of 16 through LDS, 4 x 4
// outputs per thread, fused multiply-adds.
struct GemmDims {
  int64_t M, N, K;
  int64_t sAb, sAi, sAk, sBb, sBk, sBj, sCb, sCi, sCj;
  int64_t batches, per_elem;
};
constexpr int kTM = 64, kTN = 64, kTK = 16;

template <class T>
__device__ __forceinline__ T fma_t(T a, T b, T c) {
  if constexpr (std::is_same_v<T, float>) return __builtin_fmaf(a, b, c);
  else return __builtin_fma(a, b, c);
}

template <class T>
__global__ void __launch_bounds__(kThreads) k_gemm(GemmDims d, const T* __restrict__ A, const T* __restrict__ Bm,
                                                   T* __restrict__ C, const int32_t* __restrict__ on) {
  __shared__ T As[kTK][kTM + 1];
  __shared__ T Bs[kTK][kTN + 1];
  const int64_t bt = blockIdx.z;
  if (on != nullptr && on[bt / d.per_elem] == 0) return;
  const int64_t i0 = static_cast<int64_t>(blockIdx.y) * kTM;
  const int64_t j0 = static_cast<int64_t>(blockIdx.x) * kTN;
  const T* Ab = A + bt * d.sAb;
  const T* Bb = Bm + bt * d.sBb;
  const int tid = threadIdx.x;
  const int ty = tid / 16, tx = tid % 16;
  T acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[a][c] = T(0);
  for (int64_t k0 = 0; k0 < d.K; k0 += kTK) {
    // A tile: 64 x 16; consecutive threads walk the faster-varying index of A's layout
    for (int e = tid; e < kTM * kTK; e += kThreads) {
      int ii, kk;
      if (d.sAk == 1) {
        kk = e % kTK;
        ii = e / kTK;
      } else {
        ii = e % kTM;
        kk = e / kTM;
      }
      const int64_t gi = i0 + ii, gk = k0 + kk;
      As[kk][ii] = (gi < d.M && gk < d.K) ? Ab[gi * d.sAi + gk * d.sAk] : T(0);
    }
    for (int e = tid; e < kTN * kTK; e += kThreads) {
      int jj, kk;
      if (d.sBj == 1) {
        jj = e % kTN;
        kk = e / kTN;
      } else {
        kk = e % kTK;
        jj = e / kTK;
      }
      const int64_t gj = j0 + jj, gk = k0 + kk;
      Bs[kk][jj] = (gj < d.N && gk < d.K) ? Bb[gk * d.sBk + gj * d.sBj] : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTK; ++kk) {
      T av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = As[kk][ty * 4 + a];
#pragma unroll
      for (int c = 0; c < 4; ++c) bv[c] = Bs[kk][tx * 4 + c];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[a][c] = fma_t(av[a], bv[c], acc[a][c]);
    }
    __syncthreads();
  }
  T* Cb = C + bt * d.sCb;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t gi = i0 + ty * 4 + a;
    if (gi >= d.M) continue;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int64_t gj = j0 + tx * 4 + c;
      if (gj < d.N) Cb[gi * d.sCi + gj * d.sCj] = acc[a][c];
    }
  }
}

// The float32 GEMM on the matrix cores: v_mfma_f32_16x16x4f32 (lane l supplies A(l % 16, l / 16)
// and B(l / 16, l % 16) and holds D(4 (l / 16) + v, l % 16), v < 4). A workgroup of 4 waves owns a
// 64 x 64 tile of C, each wave 32 x 32 (2 x 2 MFMA tiles); K runs in steps of 32 through LDS
// (k-major rows of 64 + 16 slots, so that the two k rows a half-wave reads fall in distinct banks).
// Every output sums its K terms in one fixed order, whatever the batch or the grid.
constexpr int kMK = 32, kMStride = kTM + 16;
typedef float mf4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(kThreads) k_gemm_mfma(GemmDims d, const float* __restrict__ A,
                                                        const float* __restrict__ Bm, float* __restrict__ C,
                                                        const int32_t* __restrict__ on) {
  __shared__ __align__(16) float As[kMK][kMStride];
  __shared__ __align__(16) float Bs[kMK][kMStride];
  const int64_t bt = blockIdx.z;
  if (on != nullptr && on[bt / d.per_elem] == 0) return;
  const int64_t i0 = static_cast<int64_t>(blockIdx.y) * kTM;
  const int64_t j0 = static_cast<int64_t>(blockIdx.x) * kTN;
  const float* Ab = A + bt * d.sAb;
  const float* Bb = Bm + bt * d.sBb;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  mf4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[a][c] = mf4{0.f, 0.f, 0.f, 0.f};
  // 16-byte loads when every tile is full and every 4-run along the contiguous index is aligned
  const bool vec = d.M % kTM == 0 && d.N % kTN == 0 && d.K % kMK == 0 &&
                   (reinterpret_cast<uintptr_t>(Ab) & 15) == 0 && (reinterpret_cast<uintptr_t>(Bb) & 15) == 0 &&
                   (d.sAk == 1 ? d.sAi % 4 == 0 : d.sAi == 1 && d.sAk % 4 == 0) &&
                   (d.sBj == 1 ? d.sBk % 4 == 0 : d.sBk == 1 && d.sBj % 4 == 0);
  for (int64_t k0 = 0; k0 < d.K; k0 += kMK) {
    if (vec) {
      // A: 4-runs along k (rows of A) or along i; B: along j or along k
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int e = tid + r * kThreads;  // 512 runs of 4 cover the 64 x 32 tile
        if (d.sAk == 1) {
          const int ii = e >> 3, kk = (e & 7) * 4;
          const float4 q = *reinterpret_cast<const float4*>(Ab + (i0 + ii) * d.sAi + (k0 + kk));
          As[kk][ii] = q.x;
          As[kk + 1][ii] = q.y;
          As[kk + 2][ii] = q.z;
          As[kk + 3][ii] = q.w;
        } else {
          const int kk = e >> 4, ii = (e & 15) * 4;
          *reinterpret_cast<float4*>(&As[kk][ii]) = *reinterpret_cast<const float4*>(Ab + (i0 + ii) + (k0 + kk) * d.sAk);
        }
        if (d.sBj == 1) {
          const int kk = e >> 4, jj = (e & 15) * 4;
          *reinterpret_cast<float4*>(&Bs[kk][jj]) = *reinterpret_cast<const float4*>(Bb + (k0 + kk) * d.sBk + (j0 + jj));
        } else {
          const int jj = e >> 3, kk = (e & 7) * 4;
          const float4 q = *reinterpret_cast<const float4*>(Bb + (k0 + kk) + (j0 + jj) * d.sBj);
          Bs[kk][jj] = q.x;
          Bs[kk + 1][jj] = q.y;
          Bs[kk + 2][jj] = q.z;
          Bs[kk + 3][jj] = q.w;
        }
      }
    } else {
      // consecutive threads walk the faster-varying index of each operand's layout
      for (int e = tid; e < kTM * kMK; e += kThreads) {
        int ii, kk;
        if (d.sAk == 1) {
          kk = e % kMK;
          ii = e / kMK;
        } else {
          ii = e % kTM;
          kk = e / kTM;
        }
        const int64_t gi = i0 + ii, gk = k0 + kk;
        As[kk][ii] = (gi < d.M && gk < d.K) ? Ab[gi * d.sAi + gk * d.sAk] : 0.f;
      }
      for (int e = tid; e < kTN * kMK; e += kThreads) {
        int jj, kk;
        if (d.sBj == 1) {
          jj = e % kTN;
          kk = e / kTN;
        } else {
          kk = e % kMK;
          jj = e / kMK;
        }
        const int64_t gj = j0 + jj, gk = k0 + kk;
        Bs[kk][jj] = (gj < d.N && gk < d.K) ? Bb[gk * d.sBk + gj * d.sBj] : 0.f;
      }
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kMK; ks += 4) {
      float av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) av[a] = As[ks + lk][wm + 16 * a + lr];
#pragma unroll
      for (int c = 0; c < 2; ++c) bv[c] = Bs[ks + lk][wn + 16 * c + lr];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bv[c], acc[a][c], 0, 0, 0);
    }
    __syncthreads();
  }
  float* Cb = C + bt * d.sCb;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gi = i0 + wm + 16 * a + 4 * lk + v;
        const int64_t gj = j0 + wn + 16 * c + lr;
        if (gi < d.M && gj < d.N) Cb[gi * d.sCi + gj * d.sCj] = acc[a][c][v];
      }
}

// The float32 GEMM as split-BF16 products on the matrix cores: every operand value v is held as
// hi = bf16(v) and lo = bf16(v - hi) (16 significant bits together), and each product as the three
// terms lo_a hi_b + hi_a lo_b + hi_a hi_b (the lo lo term, 2^-16 relative, dropped) accumulated in
// float32 by v_mfma_f32_16x16x32_bf16 (lane l supplies A(l % 16, 8 (l / 16) + e) and
// B(8 (l / 16) + e, l % 16), e < 8, and holds D(4 (l / 16) + v, l % 16), v < 4). The tiles are those
// of k_gemm_mfma (64 x 64 per workgroup, 32 x 32 per wave, K in steps of 32), staged in LDS as bf16
// rows along k (16-byte operand reads).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
constexpr int kBK = 32, kBRow = kBK + 8;

__device__ __forceinline__ void split_bf16(float v, __bf16* hi, __bf16* lo) {
  const __bf16 h = static_cast<__bf16>(v);
  *hi = h;
  *lo = static_cast<__bf16>(v - static_cast<float>(h));
}
__device__ __forceinline__ void split4(const float4& q, bf16x4& h, bf16x4& l) {
  __bf16 hh[4], ll[4];
  split_bf16(q.x, &hh[0], &ll[0]);
  split_bf16(q.y, &hh[1], &ll[1]);
  split_bf16(q.z, &hh[2], &ll[2]);
  split_bf16(q.w, &hh[3], &ll[3]);
  h = bf16x4{hh[0], hh[1], hh[2], hh[3]};
  l = bf16x4{ll[0], ll[1], ll[2], ll[3]};
}

__global__ void __launch_bounds__(kThreads) k_gemm_bf16x3(GemmDims d, const float* __restrict__ A,
                                                          const float* __restrict__ Bm, float* __restrict__ C,
                                                          const int32_t* __restrict__ on) {
  __shared__ __align__(16) __bf16 Ah[kTM][kBRow];
  __shared__ __align__(16) __bf16 Al[kTM][kBRow];
  __shared__ __align__(16) __bf16 Bh[kTN][kBRow];
  __shared__ __align__(16) __bf16 Bl[kTN][kBRow];
  const int64_t bt = blockIdx.z;
  if (on != nullptr && on[bt / d.per_elem] == 0) return;
  const int64_t i0 = static_cast<int64_t>(blockIdx.y) * kTM;
  const int64_t j0 = static_cast<int64_t>(blockIdx.x) * kTN;
  const float* Ab = A + bt * d.sAb;
  const float* Bb = Bm + bt * d.sBb;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  mf4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[a][c] = mf4{0.f, 0.f, 0.f, 0.f};
  const bool vec = d.M % kTM == 0 && d.N % kTN == 0 && d.K % kBK == 0 &&
                   (reinterpret_cast<uintptr_t>(Ab) & 15) == 0 && (reinterpret_cast<uintptr_t>(Bb) & 15) == 0 &&
                   (d.sAk == 1 ? d.sAi % 4 == 0 : d.sAi == 1 && d.sAk % 4 == 0) &&
                   (d.sBk == 1 ? d.sBj % 4 == 0 : d.sBj == 1 && d.sBk % 4 == 0);
  for (int64_t k0 = 0; k0 < d.K; k0 += kBK) {
    if (vec) {
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int e = tid + r * kThreads;  // 512 runs of 4 cover a 64 x 32 tile
        if (d.sAk == 1) {
          const int ii = e >> 3, kk = (e & 7) * 4;
          const float4 q = *reinterpret_cast<const float4*>(Ab + (i0 + ii) * d.sAi + (k0 + kk));
          bf16x4 h, l;
          split4(q, h, l);
          *reinterpret_cast<bf16x4*>(&Ah[ii][kk]) = h;
          *reinterpret_cast<bf16x4*>(&Al[ii][kk]) = l;
        } else {
          const int kk = e >> 4, ii = (e & 15) * 4;
          const float4 q = *reinterpret_cast<const float4*>(Ab + (i0 + ii) + (k0 + kk) * d.sAk);
          split_bf16(q.x, &Ah[ii][kk], &Al[ii][kk]);
          split_bf16(q.y, &Ah[ii + 1][kk], &Al[ii + 1][kk]);
          split_bf16(q.z, &Ah[ii + 2][kk], &Al[ii + 2][kk]);
          split_bf16(q.w, &Ah[ii + 3][kk], &Al[ii + 3][kk]);
        }
        if (d.sBk == 1) {
          const int jj = e >> 3, kk = (e & 7) * 4;
          const float4 q = *reinterpret_cast<const float4*>(Bb + (k0 + kk) + (j0 + jj) * d.sBj);
          bf16x4 h, l;
          split4(q, h, l);
          *reinterpret_cast<bf16x4*>(&Bh[jj][kk]) = h;
          *reinterpret_cast<bf16x4*>(&Bl[jj][kk]) = l;
        } else {
          const int kk = e >> 4, jj = (e & 15) * 4;
          const float4 q = *reinterpret_cast<const float4*>(Bb + (k0 + kk) * d.sBk + (j0 + jj));
          split_bf16(q.x, &Bh[jj][kk], &Bl[jj][kk]);
          split_bf16(q.y, &Bh[jj + 1][kk], &Bl[jj + 1][kk]);
          split_bf16(q.z, &Bh[jj + 2][kk], &Bl[jj + 2][kk]);
          split_bf16(q.w, &Bh[jj + 3][kk], &Bl[jj + 3][kk]);
        }
      }
    } else {
      for (int e = tid; e < kTM * kBK; e += kThreads) {
        int ii, kk;
        if (d.sAk == 1) {
          kk = e % kBK;
          ii = e / kBK;
        } else {
          ii = e % kTM;
          kk = e / kTM;
        }
        const int64_t gi = i0 + ii, gk = k0 + kk;
        split_bf16((gi < d.M && gk < d.K) ? Ab[gi * d.sAi + gk * d.sAk] : 0.f, &Ah[ii][kk], &Al[ii][kk]);
      }
      for (int e = tid; e < kTN * kBK; e += kThreads) {
        int jj, kk;
        if (d.sBj == 1) {
          jj = e % kTN;
          kk = e / kTN;
        } else {
          kk = e % kBK;
          jj = e / kBK;
        }
        const int64_t gj = j0 + jj, gk = k0 + kk;
        split_bf16((gj < d.N && gk < d.K) ? Bb[gk * d.sBk + gj * d.sBj] : 0.f, &Bh[jj][kk], &Bl[jj][kk]);
      }
    }
    __syncthreads();
    bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      ah[a] = *reinterpret_cast<const bf16x8*>(&Ah[wm + 16 * a + lr][lk]);
      al[a] = *reinterpret_cast<const bf16x8*>(&Al[wm + 16 * a + lr][lk]);
      bh[a] = *reinterpret_cast<const bf16x8*>(&Bh[wn + 16 * a + lr][lk]);
      bl[a] = *reinterpret_cast<const bf16x8*>(&Bl[wn + 16 * a + lr][lk]);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[c], acc[a][c], 0, 0, 0);
        acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[c], acc[a][c], 0, 0, 0);
        acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[c], acc[a][c], 0, 0, 0);
      }
    __syncthreads();
  }
  float* Cb = C + bt * d.sCb;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gi = i0 + wm + 16 * a + (lane >> 4) * 4 + v;
        const int64_t gj = j0 + wn + 16 * c + lr;
        if (gi < d.M && gj < d.N) Cb[gi * d.sCi + gj * d.sCj] = acc[a][c][v];
      }
}


// ---- the split-BF16 axis contraction (dct = "bf16x3") -------------------------------------------------
// The same three-term products as k_gemm_bf16x3, laid out for one axis pass of a DCT: the n x n
// matrix P (P[j][k], output j) is split into bf16 hi / lo once (split_table) and read as such; a
// workgroup owns 64 lines and up to 256 outputs of them, so the data is read from HBM once and split
// once per 256 outputs (the generic kernel re-read and re-split both operands for every 64 x 64
// output tile: 3.4x the algorithmic bytes, 11 % of the matrix cores' cycles busy). K runs in steps of
// 32 through LDS, rows along k (16-byte operand reads), bf16 pairs hi / lo in separate arrays.
//   contiguous axis: C[l][j] = sum_k X[l][k] P[j][k]: the MFMA rows are the 64 lines, the columns the
//     outputs (each wave 64 of them: 4 x 4 tiles of 16 x 16);
//   strided axis:    C[j][i] = sum_k P[j][k] X[k][i] per (element, slab): rows the outputs (each wave
//     64), columns 64 adjacent lines i, so that the stores run along i.
// Every output sums its K terms in one fixed order (k0 ascending, the three products lo.hi, hi.lo,
// hi.hi per step, as k_gemm_bf16x3), whatever the batch or the grid.
constexpr int kAxK = 32, kAxRow = kAxK + 8, kAxLines = 64, kAxOut = 256;
constexpr int kAxLT = kAxLines / 16;  // line tiles of 16 per wave

struct AxisGemm {
  int64_t lines;     // contiguous: lines per element (count / n); strided: inner (adjacent lines)
  int64_t slabs;     // strided: slabs per element (outer); contiguous: 1
  int32_t n;         // K = the outputs = the axis length (a multiple of 32)
  int64_t count;     // samples per element
  int64_t groups;    // line groups of kAxLines per slab
  int32_t oblocks;   // output blocks of 256
};

template <bool STRIDED>
__global__ void __launch_bounds__(kThreads) k_axis_bf16x3(AxisGemm a, const __bf16* __restrict__ Ph,
                                                          const __bf16* __restrict__ Pl,
                                                          const float* __restrict__ X, float* __restrict__ C,
                                                          const int32_t* __restrict__ on) {
  __shared__ __align__(16) __bf16 Dh[kAxLines][kAxRow];  // the data: line i, k
  __shared__ __align__(16) __bf16 Dl[kAxLines][kAxRow];
  __shared__ __align__(16) __bf16 Mh[kAxOut][kAxRow];    // the matrix: output j, k
  __shared__ __align__(16) __bf16 Ml[kAxOut][kAxRow];
  // blockIdx.x: (slab, line group); blockIdx.y: output block; blockIdx.z: element
  const int64_t b = blockIdx.z;
  if (on != nullptr && on[b] == 0) return;
  const int64_t slab = blockIdx.x / a.groups;
  const int64_t l0 = (blockIdx.x - slab * a.groups) * kAxLines;
  const int32_t j0 = static_cast<int32_t>(blockIdx.y) * kAxOut;
  const int32_t n = a.n;
  const int32_t nout = min(kAxOut, n - j0);
  const int64_t base = b * a.count + slab * static_cast<int64_t>(n) * (STRIDED ? a.lines : 1);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  // contiguous: acc[line tile][output tile]; strided: acc[output tile][line tile]
  mf4 acc[STRIDED ? 4 : kAxLT][STRIDED ? kAxLT : 4];
#pragma unroll
  for (int p = 0; p < (STRIDED ? 4 : kAxLT); ++p)
#pragma unroll
    for (int q = 0; q < (STRIDED ? kAxLT : 4); ++q) acc[p][q] = mf4{0.f, 0.f, 0.f, 0.f};
  // the global reads of step k0 go to registers one step ahead (their latency overlaps the
  // previous step's products); the data tile: kAxLines lines x 32 k (lines past the end: zeros);
  // the matrix tile: nout outputs x 32 k of the split table (16-byte reads; rows past nout: zeros)
  constexpr int kXR = kAxLines / 32;  // data runs of 4 per thread
  float4 px[kXR];
  bf16x8 pmh[4], pml[4];
  const auto fetch = [&](int32_t k0) {
#pragma unroll
    for (int r = 0; r < kXR; ++r) {
      const int e = tid + r * kThreads;
      px[r] = float4{0.f, 0.f, 0.f, 0.f};
      if constexpr (!STRIDED) {
        const int ii = e >> 3, kk = (e & 7) * 4;
        if (l0 + ii < a.lines) px[r] = *reinterpret_cast<const float4*>(X + base + (l0 + ii) * n + k0 + kk);
      } else {
        const int kk = e / (kAxLines / 4), ii = (e % (kAxLines / 4)) * 4;
        if (l0 + ii < a.lines) px[r] = *reinterpret_cast<const float4*>(X + base + (k0 + kk) * a.lines + l0 + ii);
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = tid + r * kThreads;  // 1024 runs of 8 bf16
      const int jj = e >> 2, kk = (e & 3) * 8;
      pmh[r] = bf16x8{};
      pml[r] = bf16x8{};
      if (jj < nout) {
        const int64_t at = static_cast<int64_t>(j0 + jj) * n + k0 + kk;
        pmh[r] = *reinterpret_cast<const bf16x8*>(Ph + at);
        pml[r] = *reinterpret_cast<const bf16x8*>(Pl + at);
      }
    }
  };
  fetch(0);
  for (int32_t k0 = 0; k0 < n; k0 += kAxK) {
    // the fetched tiles into LDS, the data split into hi / lo on the way
#pragma unroll
    for (int r = 0; r < kXR; ++r) {
      const int e = tid + r * kThreads;
      const float4 v = px[r];
      if constexpr (!STRIDED) {
        const int ii = e >> 3, kk = (e & 7) * 4;
        bf16x4 h, l;
        split4(v, h, l);
        *reinterpret_cast<bf16x4*>(&Dh[ii][kk]) = h;
        *reinterpret_cast<bf16x4*>(&Dl[ii][kk]) = l;
      } else {
        const int kk = e / (kAxLines / 4), ii = (e % (kAxLines / 4)) * 4;
        split_bf16(v.x, &Dh[ii][kk], &Dl[ii][kk]);
        split_bf16(v.y, &Dh[ii + 1][kk], &Dl[ii + 1][kk]);
        split_bf16(v.z, &Dh[ii + 2][kk], &Dl[ii + 2][kk]);
        split_bf16(v.w, &Dh[ii + 3][kk], &Dl[ii + 3][kk]);
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = tid + r * kThreads;
      const int jj = e >> 2, kk = (e & 3) * 8;
      *reinterpret_cast<bf16x8*>(&Mh[jj][kk]) = pmh[r];
      *reinterpret_cast<bf16x8*>(&Ml[jj][kk]) = pml[r];
    }
    __syncthreads();
    if (k0 + kAxK < n) fetch(k0 + kAxK);
    // the wave's 64 outputs (4 tiles) and the data's kAxLT line tiles: contiguous, A = the data,
    // B = the outputs; strided, A = the outputs, B = the data
    bf16x8 dh[kAxLT], dl[kAxLT], mh[4], ml[4];
#pragma unroll
    for (int p = 0; p < kAxLT; ++p) {
      dh[p] = *reinterpret_cast<const bf16x8*>(&Dh[16 * p + lr][lk]);
      dl[p] = *reinterpret_cast<const bf16x8*>(&Dl[16 * p + lr][lk]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      mh[q] = *reinterpret_cast<const bf16x8*>(&Mh[64 * wave + 16 * q + lr][lk]);
      ml[q] = *reinterpret_cast<const bf16x8*>(&Ml[64 * wave + 16 * q + lr][lk]);
    }
    const bf16x8* ah = STRIDED ? mh : dh;
    const bf16x8* al = STRIDED ? ml : dl;
    const bf16x8* bh = STRIDED ? dh : mh;
    const bf16x8* bl = STRIDED ? dl : ml;
#pragma unroll
    for (int p = 0; p < (STRIDED ? 4 : kAxLT); ++p)
#pragma unroll
      for (int q = 0; q < (STRIDED ? kAxLT : 4); ++q) {
        acc[p][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[p], bh[q], acc[p][q], 0, 0, 0);
        acc[p][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[p], bl[q], acc[p][q], 0, 0, 0);
        acc[p][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[p], bh[q], acc[p][q], 0, 0, 0);
      }
    __syncthreads();
  }
  // D(4 (lane / 16) + v, lane % 16) of tile (p, q)
#pragma unroll
  for (int p = 0; p < (STRIDED ? 4 : kAxLT); ++p)
#pragma unroll
    for (int q = 0; q < (STRIDED ? kAxLT : 4); ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * p + 4 * (lane >> 4) + v;
        const int col = 16 * q + lr;
        if constexpr (!STRIDED) {
          const int64_t l = l0 + row;
          const int32_t j = 64 * wave + col;
          if (l < a.lines && j < nout) C[base + l * n + j0 + j] = acc[p][q][v];
        } else {
          const int32_t j = 64 * wave + row;
          const int64_t i = l0 + col;
          if (i < a.lines && j < nout) C[base + static_cast<int64_t>(j0 + j) * a.lines + i] = acc[p][q][v];
        }
      }
}

// The split table of a DCT matrix (n x n, row-major, rows the outputs): hi = bf16(v), lo = bf16(v -
// hi), built once per matrix on a stream of its own (the matrices live as long as the process).
struct SplitTab {
  __bf16* hi = nullptr;
  __bf16* lo = nullptr;
};
const SplitTab& split_table(const float* M, int32_t n) {
  static std::mutex mu;
  static std::map<const float*, SplitTab> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(M);
  if (it != cache.end()) return it->second;
  const size_t nn = static_cast<size_t>(n) * n;
  SplitTab t;
  EDA_HIP(hipMalloc(&t.hi, nn * sizeof(__bf16)));
  EDA_HIP(hipMalloc(&t.lo, nn * sizeof(__bf16)));
  __bf16* hi = t.hi;
  __bf16* lo = t.lo;
  hipStream_t bs = nullptr;
  EDA_HIP(hipStreamCreateWithFlags(&bs, hipStreamNonBlocking));
  flat(bs, static_cast<int64_t>(nn), [=] __device__(int64_t e) { split_bf16(M[e], hi + e, lo + e); });
  const hipError_t built = hipStreamSynchronize(bs);
  (void)hipStreamDestroy(bs);
  EDA_HIP(built);
  return cache.emplace(M, t).first->second;
}

// The contraction takes float32 axes whose length is a multiple of 32 (and, strided, whose line
// count is a multiple of 4 for the 16-byte reads); others keep the generic split-BF16 GEMM.
bool axis_bf16_fits(const Geom& g, int a) {
  const int64_t n = g.n[a];
  const int64_t inner = g.st[a];
  return n % kAxK == 0 && (inner == 1 || inner % 4 == 0);
}
void axis_bf16x3(hipStream_t s, const Geom& g, int a, const float* M, const float* in, float* out,
                 const int32_t* on) {
  const int32_t n = static_cast<int32_t>(g.n[a]);
  const int64_t inner = g.st[a];
  const SplitTab& t = split_table(M, n);
  AxisGemm d{};
  d.n = n;
  d.count = g.count;
  d.lines = inner == 1 ? g.count / n : inner;
  d.slabs = inner == 1 ? 1 : g.count / (static_cast<int64_t>(n) * inner);
  d.groups = (d.lines + kAxLines - 1) / kAxLines;
  d.oblocks = (n + kAxOut - 1) / kAxOut;
  const dim3 grid(static_cast<unsigned>(d.slabs * d.groups), static_cast<unsigned>(d.oblocks),
                  static_cast<unsigned>(g.batch));
  if (inner == 1) EDA_GO(k_axis_bf16x3<false>, grid, dim3(kThreads), 0, s, d, t.hi, t.lo, in, out, on);
  else EDA_GO(k_axis_bf16x3<true>, grid, dim3(kThreads), 0, s, d, t.hi, t.lo, in, out, on);
  EDA_LAUNCH("k_axis_bf16x3");
}

// The DCT contraction mode of the calling thread: 0 the exact float32 forms (the LDS FFT for
// power-of-two axes, the float32 MFMA GEMM for the others), 1 every axis as a split-BF16 GEMM.
thread_local int t_dct_mode = 0;

template <class T>
void gemm(hipStream_t s, const GemmDims& d, const T* A, const T* B, T* C, const int32_t* on) {
  const dim3 grid(static_cast<unsigned>((d.N + kTN - 1) / kTN), static_cast<unsigned>((d.M + kTM - 1) / kTM),
                  static_cast<unsigned>(d.batches));
  if constexpr (std::is_same_v<T, float>) {
    if (t_dct_mode == 1) {
      EDA_GO(k_gemm_bf16x3, grid, dim3(kThreads), 0, s, d, A, B, C, on);
      EDA_LAUNCH("k_gemm_bf16x3");
      return;
    }
    EDA_GO(k_gemm_mfma, grid, dim3(kThreads), 0, s, d, A, B, C, on);
    EDA_LAUNCH("k_gemm_mfma");
  } else {
    EDA_GO(k_gemm<T>, grid, dim3(kThreads), 0, s, d, A, B, C, on);
    EDA_LAUNCH("k_gemm");
  }
}

// One axis of a separable transform: out = M applied along axis a of every running element
// (M row-major (out index, in index), n x n).
template <class T>
void axis_gemm(hipStream_t s, const Geom& g, int a, const T* M, const T* in, T* out, const int32_t* on) {
  if constexpr (std::is_same_v<T, float>) {
    if (t_dct_mode == 1 && axis_bf16_fits(g, a)) {
      axis_bf16x3(s, g, a, M, in, out, on);
      return;
    }
  }
  const int64_t n = g.n[a];
  const int64_t inner = g.st[a];
  const int64_t outer = g.count / (n * inner);
  GemmDims d{};
  if (inner == 1) {
    // per element: C (outer x n) = X (outer x n) M^T
    d.M = outer;
    d.N = n;
    d.K = n;
    d.sAb = g.count, d.sAi = n, d.sAk = 1;
    d.sBb = 0, d.sBk = 1, d.sBj = n;
    d.sCb = g.count, d.sCi = n, d.sCj = 1;
    d.batches = g.batch;
    d.per_elem = 1;
    gemm<T>(s, d, in, M, out, on);
  } else {
    // per (element, outer block): C (n x inner) = M (n x n) X (n x inner)
    d.M = n;
    d.N = inner;
    d.K = n;
    d.sAb = 0, d.sAi = n, d.sAk = 1;
    d.sBb = n * inner, d.sBk = inner, d.sBj = 1;
    d.sCb = n * inner, d.sCi = inner, d.sCj = 1;
    d.batches = g.batch * outer;
    d.per_elem = outer;
    gemm<T>(s, d, M, in, out, on);
  }
}
// The pass timer's DCT passes (kind 2) count one read and one write of the field per axis.
template <class T>
void axis_transform(hipStream_t s, const Geom& g, int a, const T* M, const T* in, T* out, const int32_t* on) {
  const bool timed = timer_gemm_on();
  hipEvent_t t0 = nullptr;
  if (timed) timer_note(s, &t0);
  axis_gemm<T>(s, g, a, M, in, out, on);
  if (timed) timer_done(s, t0, 2.0 * static_cast<double>(g.batch) * g.count * sizeof(T), kTagDct);
}

// ---- the fast transform of power-of-two axes --------------------------------------------------------
// An axis of n = 2^m samples (16 <= n <= kFftTile) is transformed in LDS through Makhoul's
// reordering and an n-point complex FFT, two lines at a time (one in the real part, one in the
// imaginary part): memory-bound, one read and one write of the field per axis.
//   forward:  v(m) = f(2m), v(n-1-m) = f(2m+1); Z = FFT(v1 + i v2); V1 = (Z_k + conj Z_{n-k}) / 2,
//             V2 = (Z_k - conj Z_{n-k}) / 2i; X_k = c_k Re(e^{-i pi k / 2n} V_k)
//   inverse:  X'_k = s_k X_k (s_0 = 1 / sqrt n, else 1 / sqrt 2n); W_k = e^{i pi k / 2n}
//             (X'_k - i X'_{n-k}) (X'_n = 0); z = sum_k (W1 + i W2)_k e^{2 pi i m k / n};
//             f(2m) = v(m), f(2m+1) = v(n-1-m) with v1 = Re z, v2 = Im z
//   sin basis on the axis: the inverse of the reversed spectrum (X_{n-k}, 0 at k = 0), times (-1)^x
// The FFT is Stockham's self-sorting form, radix 8 and one last radix 2 or 4 stage, each butterfly
// reading its R inputs n / R apart and writing its outputs Ns apart (Ns the product of the earlier
// radices), in place in LDS with a barrier between the reads and the writes. The two lines of a pair
// are two lines of one element (lines 2q and 2q + 1, a zero line past the last), so an element's
// transform depends on its own samples only.
constexpr int kFftTile = 4096;  // complex points per workgroup (float32; half for float64)
// LDS slot of point m of line pair p: one pad word per 8 points inside a line, and an odd stride
// sp = fpad(n) | 1 between pairs, so that the points one wave touches fall in distinct banks whether
// its lanes walk one line or the same point of adjacent pairs; the imaginary parts start 16 banks
// after the real parts.
__host__ __device__ __forceinline__ int32_t fpad(int32_t i) { return i + (i >> 3); }
__host__ __device__ __forceinline__ int32_t fft_sp(int32_t n) { return fpad(n) | 1; }
__host__ __device__ __forceinline__ int32_t fft_im_off(int32_t L, int32_t sp) {
  const int32_t base = L * sp;
  return base + ((16 - base % 32) + 32) % 32;
}

template <class T>
struct FftTabs {
  int32_t n = 0;
  const T* tc = nullptr;  // cos(2 pi t / n), t < n
  const T* ts = nullptr;  // sin(2 pi t / n)
  const T* pc = nullptr;  // c_k cos(pi k / 2n)
  const T* ps = nullptr;  // c_k sin(pi k / 2n)
  const T* qc = nullptr;  // cos(pi k / 2n)
  const T* qs = nullptr;  // sin(pi k / 2n)
};

// One radix-R Stockham stage over the tile (L pairs of n points), sign -1 (forward) or +1.
__device__ __forceinline__ void wave_sync_o() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// WAVE: each wave owns the pairs p = wave + 4 i of the tile (L a multiple of 4) and syncs alone;
// else the workgroup's threads share every pair.
template <class T, int R, int SIGN, bool WAVE>
__device__ __forceinline__ void fft_stage(T* re, T* im, int32_t ln, int32_t lL, int32_t lNs, int32_t sp,
                                          const FftTabs<T>& tb) {
  constexpr int IPT = (sizeof(T) == 4 ? kFftTile : kFftTile / 2) / (R * 256);  // butterflies per thread
  constexpr int LR = R == 8 ? 3 : (R == 4 ? 2 : 1);
  const int32_t n = 1 << ln;
  const int32_t lper = ln - LR;
  const int32_t per = 1 << lper;
  const int32_t total = WAVE ? per << (lL - 2) : per << lL;  // butterflies of the wave / workgroup
  const int32_t me = WAVE ? static_cast<int32_t>(threadIdx.x & 63) : static_cast<int32_t>(threadIdx.x);
  constexpr int32_t kStep = WAVE ? 64 : 256;
  const int32_t wave = static_cast<int32_t>(threadIdx.x >> 6);
  const int32_t Ns = 1 << lNs;
  const int32_t ltstep = ln - lNs - LR;  // the twiddle index step: n / (Ns R)
  const auto pair_of = [&](int32_t u) { return WAVE ? wave + 4 * (u >> lper) : u >> lper; };
  T ar[IPT][R], ai[IPT][R];
#pragma unroll
  for (int it = 0; it < IPT; ++it) {
    const int32_t u = me + it * kStep;
    if (u >= total) continue;  // (continue, not break: the loop stays unrolled, the arrays in registers)
    const int32_t p = pair_of(u);
    const int32_t j = u & (per - 1);
    const int32_t k = j & (Ns - 1);
    const int32_t base = p * sp;
    // the twiddles w^r, w = e^{SIGN 2 pi i k / (Ns R)}: w, w^2 (and w^4) from the table, the other
    // powers as their products
    T wr[R], wi[R];
    wr[0] = T(1);
    wi[0] = T(0);
    if (lNs > 0) {
      const auto tab = [&](int q, T& c, T& sn) {
        const int32_t t = ((k * q) << ltstep) & (n - 1);
        c = tb.tc[t];
        sn = SIGN < 0 ? -tb.ts[t] : tb.ts[t];
      };
      tab(1, wr[1], wi[1]);
      if constexpr (R >= 4) {
        tab(2, wr[2], wi[2]);
        wr[3] = wr[1] * wr[2] - wi[1] * wi[2];
        wi[3] = wr[1] * wi[2] + wi[1] * wr[2];
      }
      if constexpr (R == 8) {
        tab(4, wr[4], wi[4]);
#pragma unroll
        for (int q = 5; q < 8; ++q) {
          wr[q] = wr[q - 4] * wr[4] - wi[q - 4] * wi[4];
          wi[q] = wr[q - 4] * wi[4] + wi[q - 4] * wr[4];
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int32_t at = base + fpad(j + r * per);
      T xr = re[at], xi = im[at];
      if (r > 0 && lNs > 0) {
        const T yr = xr * wr[r] - xi * wi[r];
        xi = xr * wi[r] + xi * wr[r];
        xr = yr;
      }
      ar[it][r] = xr;
      ai[it][r] = xi;
    }
  }
  if (WAVE) wave_sync_o();
  else __syncthreads();
#pragma unroll
  for (int it = 0; it < IPT; ++it) {
    const int32_t u = me + it * kStep;
    if (u >= total) continue;
    const int32_t p = pair_of(u);
    const int32_t j = u & (per - 1);
    const int32_t k = j & (Ns - 1);
    T* xr = ar[it];
    T* xi = ai[it];
    // the R-point DFT with the factor e^{SIGN 2 pi i / R}; (a, b) * (SIGN i) = (-SIGN b, SIGN a)
    if constexpr (R == 2) {
      const T r0 = xr[0] + xr[1], i0 = xi[0] + xi[1];
      xr[1] = xr[0] - xr[1];
      xi[1] = xi[0] - xi[1];
      xr[0] = r0;
      xi[0] = i0;
    } else {
      // radix 4 on (0, 2, 4, 6) / 2 ... written for the general even R = 4 or 8
      constexpr int H = R / 2;
      T er[H], ei[H], orr[H], oi[H];
      if constexpr (R == 4) {
        // evens (0, 2), odds (1, 3): 2-point DFTs, then the combine with w4 = SIGN i
        er[0] = xr[0] + xr[2]; ei[0] = xi[0] + xi[2];
        er[1] = xr[0] - xr[2]; ei[1] = xi[0] - xi[2];
        orr[0] = xr[1] + xr[3]; oi[0] = xi[1] + xi[3];
        const T dr = xr[1] - xr[3], di = xi[1] - xi[3];
        orr[1] = -SIGN * di; oi[1] = SIGN * dr;
      } else {
        // evens (0, 2, 4, 6) and odds (1, 3, 5, 7) by 4-point DFTs
        const T b0r = xr[0] + xr[4], b0i = xi[0] + xi[4];
        const T b1r = xr[0] - xr[4], b1i = xi[0] - xi[4];
        const T b2r = xr[2] + xr[6], b2i = xi[2] + xi[6];
        const T d2r = xr[2] - xr[6], d2i = xi[2] - xi[6];
        const T b3r = -SIGN * d2i, b3i = SIGN * d2r;
        er[0] = b0r + b2r; ei[0] = b0i + b2i;
        er[2] = b0r - b2r; ei[2] = b0i - b2i;
        er[1] = b1r + b3r; ei[1] = b1i + b3i;
        er[3] = b1r - b3r; ei[3] = b1i - b3i;
        const T b4r = xr[1] + xr[5], b4i = xi[1] + xi[5];
        const T b5r = xr[1] - xr[5], b5i = xi[1] - xi[5];
        const T b6r = xr[3] + xr[7], b6i = xi[3] + xi[7];
        const T d6r = xr[3] - xr[7], d6i = xi[3] - xi[7];
        const T b7r = -SIGN * d6i, b7i = SIGN * d6r;
        orr[0] = b4r + b6r; oi[0] = b4i + b6i;
        orr[2] = b4r - b6r; oi[2] = b4i - b6i;
        orr[1] = b5r + b7r; oi[1] = b5i + b7i;
        orr[3] = b5r - b7r; oi[3] = b5i - b7i;
        // odds times w8^q = e^{SIGN 2 pi i q / 8}, q = 1, 2, 3
        const T h = static_cast<T>(0.70710678118654752440);
        {
          const T a = orr[1], b = oi[1];  // (a + ib)(1 + SIGN i) h
          orr[1] = (a - SIGN * b) * h;
          oi[1] = (b + SIGN * a) * h;
        }
        {
          const T a = orr[2], b = oi[2];  // (a + ib)(SIGN i)
          orr[2] = -SIGN * b;
          oi[2] = SIGN * a;
        }
        {
          const T a = orr[3], b = oi[3];  // (a + ib)(-1 + SIGN i) h
          orr[3] = (-a - SIGN * b) * h;
          oi[3] = (-b + SIGN * a) * h;
        }
      }
#pragma unroll
      for (int q = 0; q < H; ++q) {
        xr[q] = er[q] + orr[q];
        xi[q] = ei[q] + oi[q];
        xr[q + H] = er[q] - orr[q];
        xi[q + H] = ei[q] - oi[q];
      }
    }
    const int32_t dst = ((j >> lNs) << (lNs + LR)) + k;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int32_t at = p * sp + fpad(dst + r * Ns);
      re[at] = xr[r];
      im[at] = xi[r];
    }
  }
  if (WAVE) wave_sync_o();
  else __syncthreads();
}

// The stages from Ns = 2^LNS on, for a length 2^LN known at compile time (its index arithmetic folds).
template <class T, int SIGN, bool WAVE, int LN, int LNS>
__device__ __forceinline__ void fft_stages(T* re, T* im, int32_t lL, int32_t sp, const FftTabs<T>& tb) {
  if constexpr (LN - LNS >= 3) {
    fft_stage<T, 8, SIGN, WAVE>(re, im, LN, lL, LNS, sp, tb);
    fft_stages<T, SIGN, WAVE, LN, LNS + 3>(re, im, lL, sp, tb);
  } else if constexpr (LN - LNS == 2) {
    fft_stage<T, 4, SIGN, WAVE>(re, im, LN, lL, LNS, sp, tb);
  } else if constexpr (LN - LNS == 1) {
    fft_stage<T, 2, SIGN, WAVE>(re, im, LN, lL, LNS, sp, tb);
  }
}

// LN > 0: the length 2^LN at compile time; LN = 0: ln at run time.
template <class T, int SIGN, bool WAVE, int LN>
__device__ void fft_tile(T* re, T* im, int32_t ln, int32_t lL, int32_t sp, const FftTabs<T>& tb) {
  if constexpr (LN > 0) {
    fft_stages<T, SIGN, WAVE, LN, 0>(re, im, lL, sp, tb);
  } else {
    int32_t lNs = 0;
    while (ln - lNs >= 3) {
      fft_stage<T, 8, SIGN, WAVE>(re, im, ln, lL, lNs, sp, tb);
      lNs += 3;
    }
    if (ln - lNs == 2) fft_stage<T, 4, SIGN, WAVE>(re, im, ln, lL, lNs, sp, tb);
    else if (ln - lNs == 1) fft_stage<T, 2, SIGN, WAVE>(re, im, ln, lL, lNs, sp, tb);
  }
}

struct FftDims {
  int32_t n;       // axis length
  int32_t ln, lL;  // log2 n, log2 L
  int64_t inner;   // sample stride along the axis
  int64_t count;   // samples per element
  int64_t lines;   // lines per element (count / n)
  int64_t pairs;   // (lines + 1) / 2
  int32_t L;       // pairs per workgroup
  int64_t tiles;   // workgroups per element
  int64_t batch;
  int32_t xcd;     // the XCD rotation's chunk (0: none): neighbouring column tiles share cache lines
};

// The line l of element b: sample x at b count + (l / inner) n inner + x inner + l mod inner.
__device__ __forceinline__ int64_t fft_at(const FftDims& d, int64_t b, int64_t l, int32_t x) {
  const int64_t o = l / d.inner;
  return b * d.count + o * d.n * d.inner + static_cast<int64_t>(x) * d.inner + (l - o * d.inner);
}

// INV 0: the orthonormal DCT-II along the axis; INV 1: its inverse with the cos basis (odd 0) or
// the sin basis (odd 1).
template <class T, int INV, int LN>
__global__ void __launch_bounds__(256) k_dct_fft(FftDims d, FftTabs<T> tb, const T* __restrict__ in,
                                                 T* __restrict__ out, int32_t odd, const int32_t* __restrict__ on) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int32_t n = LN > 0 ? (1 << LN) : d.n;
  const int32_t L = d.L;
  const int32_t sp = fft_sp(n);
  T* re = reinterpret_cast<T*>(smem);
  const int32_t im_off = fft_im_off(L, sp);
  T* im = re + im_off;
  // the tables in LDS: the FFT twiddles, then this direction's pre / post factors
  T* tab = im + L * sp;
  FftTabs<T> tl;
  tl.n = n;
  tl.tc = tab;
  tl.ts = tab + n;
  tl.pc = tl.qc = tab + 2 * n;
  tl.ps = tl.qs = tab + 3 * n;
  {
    const T* xc = INV ? tb.qc : tb.pc;
    const T* xs = INV ? tb.qs : tb.ps;
    for (int32_t t = threadIdx.x; t < n; t += 256) {
      const T a = tb.tc[t], bsn = tb.ts[t], c = xc[t], sn = xs[t];
      tab[t] = a;
      tab[n + t] = bsn;
      tab[2 * n + t] = c;
      tab[3 * n + t] = sn;
    }
  }
  const int32_t ln = LN > 0 ? LN : d.ln, l2L = d.lL + 1;  // log2 n, log2 (2 L)
  const int64_t wg = d.xcd ? xcd_item(blockIdx.x, d.xcd) : static_cast<int64_t>(blockIdx.x);
  const int64_t b = wg / d.tiles;
  if (b >= d.batch || (on != nullptr && on[b] == 0)) return;  // uniform per workgroup
  const int64_t q0 = (wg - b * d.tiles) * L;                  // the tile's first pair
  const int32_t pts = L * n;
  // loads: 2 L lines; along the contiguous axis sample-fastest, else line-fastest (adjacent columns)
  const bool contiguous = d.inner == 1;
  // a strided axis with 2 L <= 256: each thread keeps one line of the tile (its offset computed once)
  const bool fixed_line = !contiguous && 2 * L <= 256;
  // offsets inside the element (< 2^31 samples); in and out advance to the element
  in += b * d.count;
  out += b * d.count;
  const int32_t inner = static_cast<int32_t>(d.inner);
  int32_t my_off = 0;
  if (fixed_line) {
    const int32_t l = static_cast<int32_t>(2 * q0) + static_cast<int32_t>(threadIdx.x & (2 * L - 1));
    const int32_t o = l / inner;
    my_off = o * n * inner + (l - o * inner);
  }
  const auto addr = [=](int64_t l, int32_t x) -> int32_t {
    if (contiguous) return static_cast<int32_t>(l) * n + x;
    if (fixed_line) return my_off + x * inner;
    return static_cast<int32_t>(fft_at(d, 0, l, x));
  };
  // WAVE (L a multiple of 4): wave w owns the pairs w, w + 4, ..: on the contiguous axis it loads,
  // transforms and stores them alone; on a strided axis the workgroup loads and stores the tile
  // together (adjacent columns) and the waves transform their pairs alone
  const bool wave_mode = (L & 3) == 0;
  const int32_t wave = static_cast<int32_t>(threadIdx.x >> 6);
  const int32_t lane = static_cast<int32_t>(threadIdx.x & 63);
  const bool wave_io = wave_mode && contiguous;
  // the i-th load / store item of this thread: (tile line li, sample x)
  const int32_t n_io = wave_io ? pts / 2 : 2 * pts;  // items per wave (its L / 4 pairs) or per workgroup
  const int32_t io_first = wave_io ? lane : static_cast<int32_t>(threadIdx.x);
  const int32_t io_step = wave_io ? 64 : 256;
  struct Io {
    int32_t li, x;
  };
  const auto io_item = [=](int32_t u) -> Io {
    if (wave_io) {
      const int32_t lw = u >> ln;  // the wave's line: pair wave + 4 (lw / 2), line lw mod 2
      return Io{2 * (wave + 4 * (lw >> 1)) + (lw & 1), u & (n - 1)};
    }
    if (contiguous) return Io{u >> ln, u & (n - 1)};
    return Io{u & (2 * L - 1), u >> l2L};
  };
  // the LDS slot of sample x of tile line li on the way in: Makhoul's order (forward), the spectrum
  // or the reversed spectrum (inverse; slot 0 of the reversed one is zero, written apart)
  const auto slot_in = [=](int32_t li, int32_t x) -> int32_t {
    int32_t m;
    if (INV == 0) m = (x & 1) ? n - 1 - (x >> 1) : (x >> 1);
    else m = odd ? (x == 0 ? -1 : n - x) : x;
    return m < 0 ? -1 : (li & 1 ? im_off : 0) + (li >> 1) * sp + fpad(m);
  };
  // 16-byte global accesses where the layout allows: VW consecutive samples of a line (contiguous
  // axis) or the same sample of VW adjacent lines (strided axis with inner a multiple of VW); each
  // thread's vector item u: (first line li, first sample x), the lines or samples consecutive
  constexpr int VW = 16 / static_cast<int>(sizeof(T));
  constexpr int LVW = VW == 4 ? 2 : 1;
  using V16 = std::conditional_t<sizeof(T) == 4, float4, double2>;
  const bool vec = contiguous ? n >= VW : (fixed_line && inner % VW == 0 && (2 * L) % VW == 0);
  const int32_t nv_io = n_io >> LVW;
  int32_t my_off_v = 0;
  if (vec && !contiguous) {
    const int32_t l = static_cast<int32_t>(2 * q0) + VW * static_cast<int32_t>(threadIdx.x & ((2 * L >> LVW) - 1));
    const int32_t o = l / inner;
    my_off_v = o * n * inner + (l - o * inner);
  }
  const int32_t l2Lv = l2L - LVW;
  const auto vio = [=](int32_t u) -> Io {
    if (wave_io) {
      const int32_t lw = u >> (ln - LVW);
      return Io{2 * (wave + 4 * (lw >> 1)) + (lw & 1), (u & ((n >> LVW) - 1)) << LVW};
    }
    if (contiguous) return Io{u >> (ln - LVW), (u & ((n >> LVW) - 1)) << LVW};
    return Io{(u & ((2 * L >> LVW) - 1)) << LVW, u >> l2Lv};
  };
  const auto voff = [=](const Io& io) -> int32_t {
    return contiguous ? (static_cast<int32_t>(2 * q0) + io.li) * n + io.x : my_off_v + io.x * inner;
  };
  T* lds = re;
  if (vec) {
    constexpr int FV = 4;  // 16-byte reads in flight per thread
    for (int32_t u0 = io_first; u0 < nv_io; u0 += FV * io_step) {
      V16 v[FV];
#pragma unroll
      for (int f = 0; f < FV; ++f) {
        const int32_t u = u0 + f * io_step;
        const Io io = vio(u);
        const bool ok = u < nv_io && 2 * q0 + io.li < d.lines;
        v[f] = ok ? *reinterpret_cast<const V16*>(in + voff(io)) : V16{};
      }
#pragma unroll
      for (int f = 0; f < FV; ++f) {
        const int32_t u = u0 + f * io_step;
        if (u >= nv_io) continue;
        const Io io = vio(u);
        const T e[4] = {v[f].x, v[f].y, VW == 4 ? reinterpret_cast<const T*>(&v[f])[2] : T(0),
                        VW == 4 ? reinterpret_cast<const T*>(&v[f])[3] : T(0)};
#pragma unroll
        for (int j = 0; j < VW; ++j) {
          const int32_t li = contiguous ? io.li : io.li + j;
          const int32_t x = contiguous ? io.x + j : io.x;
          const int32_t at = slot_in(li, x);
          if (at >= 0) lds[at] = e[j];
          if (INV == 1 && odd && x == 0) lds[(li & 1 ? im_off : 0) + (li >> 1) * sp] = T(0);
        }
      }
    }
  } else {
    constexpr int FU = 8;  // reads in flight per thread
    for (int32_t u0 = io_first; u0 < n_io; u0 += FU * io_step) {
      T v[FU];
#pragma unroll
      for (int f = 0; f < FU; ++f) {
        const int32_t u = u0 + f * io_step;
        const Io io = io_item(u);
        const int64_t l = 2 * q0 + io.li;
        v[f] = u < n_io && l < d.lines ? in[addr(l, io.x)] : T(0);
      }
#pragma unroll
      for (int f = 0; f < FU; ++f) {
        const int32_t u = u0 + f * io_step;
        if (u >= n_io) continue;
        const Io io = io_item(u);
        const int32_t at = slot_in(io.li, io.x);
        if (at >= 0) lds[at] = v[f];
        if (INV == 1 && odd && io.x == 0) lds[(io.li & 1 ? im_off : 0) + (io.li >> 1) * sp] = T(0);
      }
    }
  }
  __syncthreads();  // the tables, and the tile when the workgroup loaded it
  if (INV == 1) {
    // W1 + i W2 for the pair (k, n - k) of every line pair, in place; k < half, then k = half
    const int32_t half = n / 2;
    const int32_t Lo = wave_mode ? L / 4 : L;  // pairs of the wave / workgroup
    const int32_t pfirst = wave_mode ? lane : static_cast<int32_t>(threadIdx.x);
    const int32_t pstep = wave_mode ? 64 : 256;
    for (int32_t u = pfirst; u < Lo * (half + 1); u += pstep) {
      const bool top = u >= Lo * half;
      const int32_t po = top ? u - Lo * half : u >> (ln - 1);
      const int32_t p = wave_mode ? wave + 4 * po : po;
      const int32_t k = top ? half : u & (half - 1);
      const int32_t kk = k == 0 ? 0 : n - k;
      const T sk = k == 0 ? static_cast<T>(1.0 / sqrt(static_cast<double>(n))) : static_cast<T>(1.0 / sqrt(2.0 * n));
      const T sr = static_cast<T>(1.0 / sqrt(2.0 * n));
      const int32_t a0 = p * sp + fpad(k), a1 = p * sp + fpad(kk);
      const T x1 = re[a0] * sk, x2 = im[a0] * sk;                            // X'_k of the two lines
      const T y1 = k == 0 ? T(0) : re[a1] * sr, y2 = k == 0 ? T(0) : im[a1] * sr;  // X'_{n-k}
      // W_k = e^{i phi_k} (X'_k - i X'_{n-k}); W_{n-k} = e^{i phi_{n-k}} (X'_{n-k} - i X'_k)
      const T c0 = tl.qc[k], s0 = tl.qs[k];
      const T w1r = c0 * x1 + s0 * y1, w1i = s0 * x1 - c0 * y1;
      const T w2r = c0 * x2 + s0 * y2, w2i = s0 * x2 - c0 * y2;
      re[a0] = w1r - w2i;
      im[a0] = w1i + w2r;
      if (k != 0 && k != half) {
        const T c1 = tl.qc[kk], s1 = tl.qs[kk];
        const T v1r = c1 * y1 + s1 * x1, v1i = s1 * y1 - c1 * x1;
        const T v2r = c1 * y2 + s1 * x2, v2i = s1 * y2 - c1 * x2;
        re[a1] = v1r - v2i;
        im[a1] = v1i + v2r;
      }
    }
  }
  if (wave_mode) {
    wave_sync_o();
    fft_tile<T, INV ? +1 : -1, true, LN>(re, im, ln, d.lL, sp, tl);
    if (!wave_io) __syncthreads();  // the workgroup stores the tile
  } else {
    __syncthreads();
    fft_tile<T, INV ? +1 : -1, false, LN>(re, im, ln, d.lL, sp, tl);
  }
  // the output value of tile line li at sample (or frequency) x
  const auto value_out = [&](int32_t li, int32_t x) -> T {
    const int32_t p = li >> 1;
    if (INV == 1) {
      // f(2m) = v(m), f(2m+1) = v(n-1-m); the sin basis: times (-1)^x
      const int32_t m = (x & 1) ? n - 1 - (x >> 1) : (x >> 1);
      const T v = ((li & 1) ? im : re)[p * sp + fpad(m)];
      return odd && (x & 1) ? -v : v;
    }
    // X_k of the pair's first line from V1 = (Z_k + conj Z_{n-k}) / 2, of its second from
    // V2 = (Z_k - conj Z_{n-k}) / 2i
    const int32_t k = x;
    const int32_t kk = k == 0 ? 0 : n - k;
    const T zr = re[p * sp + fpad(k)], zi = im[p * sp + fpad(k)];
    const T yr = re[p * sp + fpad(kk)], yi = im[p * sp + fpad(kk)];
    T vr, vi;
    if ((li & 1) == 0) {
      vr = (zr + yr) * T(0.5);
      vi = (zi - yi) * T(0.5);
    } else {
      vr = (zi + yi) * T(0.5);
      vi = (yr - zr) * T(0.5);
    }
    return tl.pc[k] * vr + tl.ps[k] * vi;
  };
  if (vec) {
    for (int32_t u = io_first; u < nv_io; u += io_step) {
      const Io io = vio(u);
      if (2 * q0 + io.li >= d.lines) continue;
      T e[4];
#pragma unroll
      for (int j = 0; j < VW; ++j) e[j] = contiguous ? value_out(io.li, io.x + j) : value_out(io.li + j, io.x);
      V16 q;
      if constexpr (VW == 4) q = V16{e[0], e[1], e[2], e[3]};
      else q = V16{e[0], e[1]};
      *reinterpret_cast<V16*>(out + voff(io)) = q;
    }
  } else {
    for (int32_t u = io_first; u < n_io; u += io_step) {
      const Io io = io_item(u);
      const int64_t l = 2 * q0 + io.li;
      if (l >= d.lines) continue;
      out[addr(l, io.x)] = value_out(io.li, io.x);
    }
  }
}

template <class T>
FftTabs<T> build_fft_tabs(int32_t n) {
  const long double pi = 3.141592653589793238462643383279502884L;
  std::vector<T> h(6 * static_cast<size_t>(n));
  for (int64_t t = 0; t < n; ++t) {
    const long double a = 2 * pi * static_cast<long double>(t) / static_cast<long double>(n);
    const long double ph = pi * static_cast<long double>(t) / static_cast<long double>(2 * n);
    const long double ck = std::sqrt((t ? 2.0L : 1.0L) / static_cast<long double>(n));
    h[t] = static_cast<T>(std::cos(a));
    h[n + t] = static_cast<T>(std::sin(a));
    h[2 * n + t] = static_cast<T>(ck * std::cos(ph));
    h[3 * n + t] = static_cast<T>(ck * std::sin(ph));
    h[4 * n + t] = static_cast<T>(std::cos(ph));
    h[5 * n + t] = static_cast<T>(std::sin(ph));
  }
  T* dv;
  EDA_HIP(hipMalloc(&dv, h.size() * sizeof(T)));
  EDA_HIP(hipMemcpy(dv, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  FftTabs<T> tb;
  tb.n = n;
  tb.tc = dv;
  tb.ts = dv + n;
  tb.pc = dv + 2 * n;
  tb.ps = dv + 3 * n;
  tb.qc = dv + 4 * n;
  tb.qs = dv + 5 * n;
  return tb;
}

template <class T>
const FftTabs<T>& fft_tabs(int device, int32_t n) {
  static std::mutex mu;
  static std::map<std::pair<int, int32_t>, FftTabs<T>> cache;
  std::lock_guard<std::mutex> lock(mu);
  const auto key = std::make_pair(device, n);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  return cache.emplace(key, build_fft_tabs<T>(n)).first->second;
}

// Axes that take the fast transform: powers of two from 16 to the tile.
template <class T>
bool fft_axis(int32_t n) {
  const int32_t tile = sizeof(T) == 4 ? kFftTile : kFftTile / 2;
  return n >= 16 && n <= tile && (n & (n - 1)) == 0;
}

template <class T>
void fft_transform(hipStream_t s, int device, const Geom& g, int a, const T* in, T* out, int inverse, int odd,
                   const int32_t* on) {
  const int32_t n = static_cast<int32_t>(g.n[a]);
  const int32_t tile = sizeof(T) == 4 ? kFftTile : kFftTile / 2;
  FftDims d{};
  d.n = n;
  d.inner = g.st[a];
  d.count = g.count;
  d.lines = g.count / n;
  d.pairs = (d.lines + 1) / 2;
  // a power of two; pairs past the element's last are zero lines. Rows of the contiguous axis take
  // half the tile (more workgroups per CU); columns of a strided axis the whole (16 adjacent columns
  // of 512 samples: 64-byte runs)
  d.L = g.st[a] == 1 ? std::max(1, tile / n / 2) : tile / n;
  while (d.L > 1 && d.L / 2 >= d.pairs) d.L /= 2;
  d.tiles = (d.pairs + d.L - 1) / d.L;
  d.ln = 0;
  while ((1 << d.ln) < n) ++d.ln;
  d.lL = 0;
  while ((1 << d.lL) < d.L) ++d.lL;
  d.batch = g.batch;
  const FftTabs<T>& tb = fft_tabs<T>(device, n);
  const size_t lds = (static_cast<size_t>(fft_im_off(d.L, fft_sp(n))) + static_cast<size_t>(d.L) * fft_sp(n) +
                      4 * static_cast<size_t>(n)) * sizeof(T);
  const int64_t grid = d.tiles * g.batch;
  d.xcd = g.st[a] != 1 ? xcd_chunk(grid) : 0;
  const bool timed = timer_gemm_on();
  hipEvent_t t0 = nullptr;
  if (timed) timer_note(s, &t0);
  const dim3 gr(static_cast<unsigned>(grid));
  const auto go = [&](auto kern) { EDA_GO(kern, gr, dim3(256), lds, s, d, tb, in, out, odd, on); };
  // lengths 2^5 .. 2^10 with their index arithmetic folded at compile time
  switch (inverse ? 100 + d.ln : d.ln) {
    case 5: go(k_dct_fft<T, 0, 5>); break;
    case 6: go(k_dct_fft<T, 0, 6>); break;
    case 7: go(k_dct_fft<T, 0, 7>); break;
    case 8: go(k_dct_fft<T, 0, 8>); break;
    case 9: go(k_dct_fft<T, 0, 9>); break;
    case 10: go(k_dct_fft<T, 0, 10>); break;
    case 105: go(k_dct_fft<T, 1, 5>); break;
    case 106: go(k_dct_fft<T, 1, 6>); break;
    case 107: go(k_dct_fft<T, 1, 7>); break;
    case 108: go(k_dct_fft<T, 1, 8>); break;
    case 109: go(k_dct_fft<T, 1, 9>); break;
    case 110: go(k_dct_fft<T, 1, 10>); break;
    default:
      if (inverse) go(k_dct_fft<T, 1, 0>);
      else go(k_dct_fft<T, 0, 0>);
  }
  EDA_LAUNCH("k_dct_fft");
  if (timed) timer_done(s, t0, 2.0 * static_cast<double>(g.batch) * g.count * sizeof(T), kTagDct);
}

// ---- matrices -----------------------------------------------------------------------------------
template <class T>
AxisMats<T> build_mats(int32_t n) {
  const int64_t N4 = 4 * static_cast<int64_t>(n);
  std::vector<double> cs(N4), sn(N4), ck(n);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int64_t j = 0; j < N4; ++j) {
    cs[j] = static_cast<double>(std::cos(pi * static_cast<long double>(j) / static_cast<long double>(2 * n)));
    sn[j] = static_cast<double>(std::sin(pi * static_cast<long double>(j) / static_cast<long double>(2 * n)));
  }
  for (int64_t k = 0; k < n; ++k) ck[k] = static_cast<double>(std::sqrt((k ? 2.0L : 1.0L) / static_cast<long double>(n)));
  double *dcs, *dsn, *dck;
  EDA_HIP(hipMalloc(&dcs, N4 * sizeof(double)));
  EDA_HIP(hipMalloc(&dsn, N4 * sizeof(double)));
  EDA_HIP(hipMalloc(&dck, n * sizeof(double)));
  EDA_HIP(hipMemcpy(dcs, cs.data(), N4 * sizeof(double), hipMemcpyHostToDevice));
  EDA_HIP(hipMemcpy(dsn, sn.data(), N4 * sizeof(double), hipMemcpyHostToDevice));
  EDA_HIP(hipMemcpy(dck, ck.data(), n * sizeof(double), hipMemcpyHostToDevice));
  AxisMats<T> m;
  m.n = n;
  const size_t nn = static_cast<size_t>(n) * n;
  EDA_HIP(hipMalloc(&m.fwd, nn * sizeof(T)));
  EDA_HIP(hipMalloc(&m.icos, nn * sizeof(T)));
  EDA_HIP(hipMalloc(&m.isin, nn * sizeof(T)));
  T* fwd = m.fwd;
  T* icos = m.icos;
  T* isin = m.isin;
  // built on a stream of its own and waited for there: the callers' streams keep running
  hipStream_t bs = nullptr;
  EDA_HIP(hipStreamCreateWithFlags(&bs, hipStreamNonBlocking));
  flat(bs, static_cast<int64_t>(nn), [=] __device__(int64_t e) {
    const int64_t k = e / n, x = e - k * n;
    const int64_t q = (k * (2 * x + 1)) % N4;
    fwd[k * n + x] = static_cast<T>(dck[k] * dcs[q]);
    icos[x * n + k] = static_cast<T>(dck[k] * dcs[q]);
    isin[x * n + k] = static_cast<T>(dck[k] * dsn[q]);
  });
  const hipError_t built = hipStreamSynchronize(bs);
  (void)hipStreamDestroy(bs);
  EDA_HIP(built);
  EDA_HIP(hipFree(dcs));
  EDA_HIP(hipFree(dsn));
  EDA_HIP(hipFree(dck));
  return m;
}

// The pair index of R2 and T: 2D (0,0) (0,1) (1,1); 3D (0,0) (0,1) (0,2) (1,1) (1,2) (2,2).
__host__ __device__ __forceinline__ int tensor_pair(int D, int a, int b) {
  if (a > b) {
    const int t = a;
    a = b;
    b = t;
  }
  return D == 2 ? (a == 0 ? b : 2) : (a == 0 ? b : (a == 1 ? 2 + b : 5));
}

// ---- the eigen pass (reference cpp/src/method/element.cu:29-164) ---------------------------------
__device__ __forceinline__ void jacobi_rotate(double (&A)[3][3], double (&V)[3][3], int p, int q, int r, int dim) {
  const double apq = A[p][q];
  if (apq == 0.0) return;
  const double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
  const double at = fabs(theta);
  const double t = copysign(1.0, theta) / (at + sqrt(theta * theta + 1.0));
  const double c = 1.0 / sqrt(t * t + 1.0);
  const double sn = t * c;
  const double tau = sn / (1.0 + c);
  const double tapq = t * apq;
  A[p][p] = A[p][p] - tapq;
  A[q][q] = A[q][q] + tapq;
  A[p][q] = A[q][p] = 0.0;
  if (r >= 0) {
    const double gg = A[r][p], hh = A[r][q];
    const double np_ = gg - sn * (hh + gg * tau);
    const double nq_ = hh + sn * (gg - hh * tau);
    A[r][p] = A[p][r] = np_;
    A[r][q] = A[q][r] = nq_;
  }
  for (int k = 0; k < dim; ++k) {
    const double gg = V[k][p], hh = V[k][q];
    V[k][p] = gg - sn * (hh + gg * tau);
    V[k][q] = hh + sn * (gg - hh * tau);
  }
}

template <int D>
__device__ __forceinline__ bool eigen_voxel(const double* t, double* nv, double* cv) {
  double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  bool finite = true;
  int k = 0;
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int c = a; c < D; ++c, ++k) {
      finite = finite && isfinite(t[k]);
      A[a][c] = A[c][a] = t[k];
    }
  if constexpr (D == 2) {
    jacobi_rotate(A, V, 0, 1, -1, 2);
  } else {
    double amax = 0.0;
    for (int a = 0; a < 3; ++a)
      for (int c = a; c < 3; ++c) amax = fmax(amax, fabs(A[a][c]));
    if (amax > 0.0 && amax < __builtin_huge_val()) {
      const double sc = scalbn(1.0, -ilogb(amax));
      for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) A[a][c] = A[a][c] * sc;
    }
#pragma unroll 1
    for (int sweep = 0; sweep < 6; ++sweep) {
      jacobi_rotate(A, V, 0, 1, 2, 3);
      jacobi_rotate(A, V, 0, 2, 1, 3);
      jacobi_rotate(A, V, 1, 2, 0, 3);
    }
  }
  double mu[D];
#pragma unroll
  for (int a = 0; a < D; ++a) mu[a] = A[a][a];
  int top = 0;
#pragma unroll
  for (int a = 1; a < D; ++a) top = mu[a] > mu[top] ? a : top;
  double mu2 = -__builtin_huge_val();
#pragma unroll
  for (int a = 0; a < D; ++a) mu2 = (a != top && mu[a] > mu2) ? mu[a] : mu2;
  double mu1 = mu[0];
#pragma unroll
  for (int a = 0; a < D; ++a)
    if (a == top) mu1 = mu[a];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double v = V[a][0];
#pragma unroll
    for (int c = 1; c < D; ++c) v = c == top ? V[a][c] : v;
    nv[a] = v;
  }
  const double ssum = mu1 + mu2;
  double c = ssum > 0.0 ? (mu1 - mu2) / ssum : 0.0;
  *cv = c < 0.0 ? 0.0 : (c > 1.0 ? 1.0 : c);
  return finite;
}

template <class T, int D>
void eigen_pass(hipStream_t s, const Geom& g, const T* t, const int32_t* on, T* nvec, T* coh, int32_t* nonfinite) {
  constexpr int K = D * (D + 1) / 2;
  const int64_t F = g.batch * g.count;
  each(s, g, on, [=] __device__(int64_t b, int64_t i) {
    const int64_t at = b * g.count + i;
    double tv[K], nv[D], cv;
#pragma unroll
    for (int p = 0; p < K; ++p) tv[p] = static_cast<double>(t[p * F + at]);
    const bool ok = eigen_voxel<D>(tv, nv, &cv);
    if (!ok && nonfinite != nullptr) nonfinite[b] = 1;
    coh[at] = static_cast<T>(cv);
#pragma unroll
    for (int a = 0; a < D; ++a) nvec[a * F + at] = static_cast<T>(nv[a]);
  });
}

// The mean of every running element in double: 256 strided partial sums and a fixed tree (the
// value depends on the element's samples only).
template <class T>
__global__ void __launch_bounds__(kThreads) k_mean_part(const T* __restrict__ r, int64_t count,
                                                        const int32_t* __restrict__ on, double* __restrict__ part) {
  __shared__ double sh[kThreads];
  const int64_t b = blockIdx.y;
  if (on != nullptr && on[b] == 0) return;
  double acc = 0.0;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kThreads) + threadIdx.x; i < count;
       i += static_cast<int64_t>(gridDim.x) * kThreads)
    acc = acc + static_cast<double>(r[b * count + i]);
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int st = kThreads / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[b * gridDim.x + blockIdx.x] = sh[0];
}
constexpr int kMeanParts = 64;
// The mean of every running element in double: kMeanParts workgroups of 256 strided partial sums
// with a fixed tree each, then their sum in order (the value depends on the element's samples only).
template <class T>
void enqueue_mean(hipStream_t s, const Geom& g, const T* r, const int32_t* on, double* mean, double* part) {
  EDA_GO(k_mean_part<T>, dim3(kMeanParts, static_cast<unsigned>(g.batch)), dim3(kThreads), 0, s, r, g.count,
                     on, part);
  EDA_LAUNCH("k_mean_part");
  const int64_t count = g.count;
  flat(s, g.batch, [=] __device__(int64_t b) {
    if (on != nullptr && on[b] == 0) return;
    double v = 0.0;
    for (int k = 0; k < kMeanParts; ++k) v = v + part[b * kMeanParts + k];
    mean[b] = v / static_cast<double>(count);
  });
}

// The DCT index k of sample i, the frequencies u_a = k_a (1 / (2 n_a)) and rad = |u| (summed from 0
// in axis order; reference orientation.cu:101-111).
__device__ __forceinline__ double dct_freq(const Geom& g, int64_t i, int32_t* k, double* u) {
  double r2 = 0.0;
  for (int a = 0; a < g.dim; ++a) {
    k[a] = static_cast<int32_t>((i / g.st[a]) % g.n[a]);
    u[a] = static_cast<double>(k[a]) * (1.0 / (2.0 * static_cast<double>(g.n[a])));
    r2 = r2 + u[a] * u[a];
  }
  return sqrt(r2);
}

// 2 sigma^2 of the one-octave log-Gabor, sigma = log 2 / (2 sqrt(2 log 2)) (orientation.cu:67-79).
double band_two_sigma2() {
  volatile double two = 2.0;
  const double l2 = std::log(static_cast<double>(two));
  const double sg = l2 / (2.0 * std::sqrt(2.0 * l2));
  return 2.0 * (sg * sg);
}

// The component kinds of the Riesz and derivative transforms (reference orientation.cu:86-128):
// the multiplier at DCT index k (0 at the DC) and the axes whose inverse uses the sin basis:
//   0 R1_a    u_a / rad                  sin on a
//   1 R2_aa   -(u_a / rad)^2             -
//   2 R2_ab   (u_a / rad) (u_b / rad)    sin on a, b   (a < b)
//   3 D_a     -u_a                       sin on a      (d/dx_a / (2 pi))
//   4 D_b R_a u_a (u_a / rad)            -             (a == b)
//             -u_a (u_b / rad)           sin on a, b   (a < b)
// The reference inverts the odd components through a spectral reversal and the sign (-1)^(x_a);
// the sin basis is the same sum without either.
__device__ __forceinline__ double riesz_mult(int kind, int a, int b, const double* u, double rad) {
  if (!(rad > 0.0)) return 0.0;
  const double inv = 1.0 / rad;
  const double ua = u[a] * inv;
  if (kind == 0) return ua;
  if (kind == 1) return (-ua) * ua;
  if (kind == 2) return ua * (u[b] * inv);
  if (kind == 3) return -u[a];
  return a == b ? u[a] * ua : -(u[a] * (u[b] * inv));
}
int riesz_odd(int kind, int a, int b) {
  if (kind == 0 || kind == 3) return 1 << a;
  if (kind == 1 || (kind == 4 && a == b)) return 0;
  return (1 << a) | (1 << b);
}

}  // namespace

int set_dct_mode(int mode) {
  const int prev = t_dct_mode;
  t_dct_mode = mode;
  return prev;
}

template <class T>
const AxisMats<T>& axis_mats(int device, int32_t n) {
  static std::mutex mu;
  static std::map<std::pair<int, int32_t>, AxisMats<T>> cache;
  std::lock_guard<std::mutex> lock(mu);
  const auto key = std::make_pair(device, n);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  return cache.emplace(key, build_mats<T>(n)).first->second;
}

template <class T>
void enqueue_dct(hipStream_t s, int device, const Geom& g, const T* in, T* out, T* tmp, int inverse, int odd) {
  // the axes from last to first, ping-ponging between tmp and out so that the last pass lands in out;
  // power-of-two axes through the fast transform, the others through their dense matrix
  const int D = g.dim;
  const T* src = in;
  for (int k = 0; k < D; ++k) {
    const int a = D - 1 - k;
    T* dst = ((D - 1 - k) % 2 == 0) ? out : tmp;
    if (fft_axis<T>(static_cast<int32_t>(g.n[a])) && !(std::is_same_v<T, float> && t_dct_mode == 1)) {
      fft_transform<T>(s, device, g, a, src, dst, inverse, (odd >> a) & 1, nullptr);
    } else {
      const AxisMats<T>& m = axis_mats<T>(device, g.n[a]);
      const T* M = !inverse ? m.fwd : (((odd >> a) & 1) ? m.isin : m.icos);
      axis_transform<T>(s, g, a, M, src, dst, nullptr);
    }
    src = dst;
  }
}

template <class T>
void enqueue_eigen(hipStream_t s, const Geom& g, const T* t, const int32_t* on, T* nvec, T* coh, int32_t* nonfinite) {
  if (g.dim == 2) eigen_pass<T, 2>(s, g, t, on, nvec, coh, nonfinite);
  else if (g.dim == 3) eigen_pass<T, 3>(s, g, t, on, nvec, coh, nonfinite);
}

// Per voxel the structure tensor T = R1 R1^T + R2 R2^T in T's arithmetic (pair order), then its
// eigen pass into n and c, in one kernel (the tensor stays in registers; nvec may be R1's memory:
// each voxel reads its R1 before it writes its n).
template <class T, int D>
void tensor_eigen_d(hipStream_t s, const Geom& g, const T* R1, const T* R2, const int32_t* on, T* nvec, T* coh,
                    int32_t* nonfinite) {
  constexpr int K = D * (D + 1) / 2;
  const int64_t F = g.batch * g.count;
  each(s, g, on, [=] __device__(int64_t b, int64_t i) {
    const int64_t at = b * g.count + i;
    T r1[D], r2[K];
#pragma unroll
    for (int a = 0; a < D; ++a) r1[a] = R1[a * F + at];
#pragma unroll
    for (int p = 0; p < K; ++p) r2[p] = R2[p * F + at];
    double tv[K];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int c = a; c < D; ++c) {
        T v = r1[a] * r1[c];
#pragma unroll
        for (int q = 0; q < D; ++q) v = v + r2[tensor_pair(D, a, q)] * r2[tensor_pair(D, q, c)];
        tv[tensor_pair(D, a, c)] = static_cast<double>(v);
      }
    double nv[D], cv;
    const bool ok = eigen_voxel<D>(tv, nv, &cv);
    if (!ok && nonfinite != nullptr) nonfinite[b] = 1;
    coh[at] = static_cast<T>(cv);
#pragma unroll
    for (int a = 0; a < D; ++a) nvec[a * F + at] = static_cast<T>(nv[a]);
  });
}
template <class T>
void tensor_eigen(hipStream_t s, const Geom& g, const T* R1, const T* R2, const int32_t* on, T* nvec, T* coh,
                  int32_t* nonfinite) {
  if (g.dim == 2) tensor_eigen_d<T, 2>(s, g, R1, R2, on, nvec, coh, nonfinite);
  else if (g.dim == 3) tensor_eigen_d<T, 3>(s, g, R1, R2, on, nvec, coh, nonfinite);
}

// Y = T(m X) for a kind of riesz_mult: the multiplier depends on the sample's position only, so
// each thread computes it once for a position and applies it to that position of every running
// element (kU elements' reads in flight at a time).
template <class T>
__global__ void __launch_bounds__(kThreads) k_riesz_mult(Geom g, const T* __restrict__ X, T* __restrict__ Y, int kind,
                                                         int a, int bb, const int32_t* __restrict__ on) {
  constexpr int kU = 8;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kThreads) + threadIdx.x; i < g.count;
       i += static_cast<int64_t>(gridDim.x) * kThreads) {
    int32_t k[kMaxDim];
    double u[kMaxDim];
    const double rad = dct_freq(g, i, k, u);
    const double m = riesz_mult(kind, a, bb, u, rad);
    for (int64_t b0 = 0; b0 < g.batch; b0 += kU) {
      T x[kU];
#pragma unroll
      for (int j = 0; j < kU; ++j) {
        const int64_t b = b0 + j;
        x[j] = b < g.batch && (on == nullptr || on[b] != 0) ? X[b * g.count + i] : T(0);
      }
#pragma unroll
      for (int j = 0; j < kU; ++j) {
        const int64_t b = b0 + j;
        if (b < g.batch && (on == nullptr || on[b] != 0)) Y[b * g.count + i] = static_cast<T>(m * static_cast<double>(x[j]));
      }
    }
  }
}

// Component ci of the orientation's D + D (D + 1) / 2: ci < D is R1_ci (kind 0); then the pairs
// (a, bb), a <= bb, in pair order: R2_aa (kind 1), R2_ab (kind 2).
inline void riesz_component_of(int D, int ci, int* a, int* bb, int* kind) {
  int aa = ci, b2 = ci;
  if (ci >= D) {
    int p = ci - D;
    for (aa = 0; aa < D; ++aa) {
      if (p < D - aa) break;
      p -= D - aa;
    }
    b2 = aa + p;
  }
  *a = aa;
  *bb = b2;
  *kind = ci < D ? 0 : (aa == b2 ? 1 : 2);
}

// One component of X (a kind of riesz_mult) into dst: Y = T(m X), then the inverse transform with
// the sin basis on the component's odd axes (T1: the transform's scratch field).
template <class T>
void riesz_component(hipStream_t s, int device, const Geom& g, const T* X, int kind, int a, int bb, T* Y, T* T1,
                     T* dst, const int32_t* on) {
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((g.count + kThreads - 1) / kThreads, 65535));
  EDA_GO(k_riesz_mult<T>, dim3(static_cast<unsigned>(blocks)), dim3(kThreads), 0, s, g, X, Y, kind, a, bb, on);
  EDA_LAUNCH("k_riesz_mult");
  enqueue_dct<T>(s, device, g, Y, dst, T1, 1, riesz_odd(kind, a, bb));
}

// The Riesz components of the band of window w[b] around 1 / w[b] (reference
// orientation.cu:86-360): R1 (D fields) and R2 (D (D + 1) / 2 fields, pair order) of r - mean(r).
// work: 3 fields (X, Y, the transform's scratch).
template <class T>
void riesz_fields(hipStream_t s, int device, const Geom& g, const T* r, const int32_t* w, const int32_t* on, T* R1,
                  T* R2, T* work, double* dscratch) {
  const int D = g.dim;
  const int64_t F = g.batch * g.count;
  const int K = D * (D + 1) / 2;
  T* X = work;
  T* Y = work + F;
  T* T1 = work + 2 * F;
  double* mean = dscratch;
  enqueue_mean<T>(s, g, r, on, mean, mean + g.batch);
  each(s, g, on, [=] __device__(int64_t b, int64_t i) {
    const int64_t at = b * g.count + i;
    Y[at] = static_cast<T>(static_cast<double>(r[at]) - mean[b]);
  });
  enqueue_dct<T>(s, device, g, Y, X, T1, 0, 0);
  // the band G(rad) = exp(-(log(rad / f0))^2 / (2 sigma^2)), f0 = 1 / w; G(0) = 0
  const double two_sigma2 = band_two_sigma2();
  each(s, g, on, [=] __device__(int64_t b, int64_t i) {
    int32_t k[kMaxDim];
    double u[kMaxDim];
    const double rad = dct_freq(g, i, k, u);
    double G = 0.0;
    if (rad > 0.0) {
      const double f0 = 1.0 / (1.0 * static_cast<double>(w[b]));
      const double l = log(rad / f0);
      G = exp(-(l * l) / two_sigma2);
    }
    const int64_t at = b * g.count + i;
    X[at] = static_cast<T>(static_cast<double>(X[at]) * G);
  });
  // the components: kind 0 R1_a (sin on a), 1 R2_aa, 2 R2_ab (sin on a and b)
  for (int ci = 0; ci < D + K; ++ci) {
    int a, bb, kind;
    riesz_component_of(D, ci, &a, &bb, &kind);
    riesz_component<T>(s, device, g, X, kind, a, bb, Y, T1, kind == 0 ? R1 + a * F : R2 + tensor_pair(D, a, bb) * F,
                       on);
  }
}

template <class T>
void enqueue_orientation(hipStream_t s, int device, const Geom& g, const T* r, const int32_t* w, const int32_t* on,
                         T* nvec, T* coh, T* work, int32_t* nonfinite, double* dscratch) {
  const int D = g.dim;
  if (D < 2) return;
  const int64_t F = g.batch * g.count;
  T* R2 = work + 3 * F;
  T* R1 = nvec;  // n is written last, over R1
  riesz_fields<T>(s, device, g, r, w, on, R1, R2, work, dscratch);
  tensor_eigen<T>(s, g, R1, R2, on, nvec, coh, nonfinite);
}

// ---- the orientation, transposed (the adjoint's path from n and c back to r) ----------------------
// Per voxel: the structure tensor T = R1 R1^T + R2 R2^T as the forward forms it, its eigen pairs
// (the forward's Jacobi rotations in double; the 3-D scaling undone), then, for the gradients gn of
// n (the top eigenvector, with the forward's sign) and gc of c = clip((mu1 - mu2) / (mu1 + mu2), 0,
// 1): dL/dT = gmu1 v1 v1^T + gmu2 v2 v2^T + sum_{j != 1} (v_j . gn) / (mu1 - mu_j) sym(v_j v1^T) (the
// first-order perturbation of a symmetric eigen problem; a term whose eigenvalue gap is 0 is
// dropped), and through T's products the gradients of R1 and R2.
template <int D>
__device__ __forceinline__ bool eigen_pairs(const double* t, double (&mu)[3], double (&V)[3][3]) {
  double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) V[a][c] = a == c ? 1.0 : 0.0;
  bool finite = true;
  int k = 0;
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int c = a; c < D; ++c, ++k) {
      finite = finite && isfinite(t[k]);
      A[a][c] = A[c][a] = t[k];
    }
  double sc = 1.0;
  if constexpr (D == 2) {
    jacobi_rotate(A, V, 0, 1, -1, 2);
  } else {
    double amax = 0.0;
    for (int a = 0; a < 3; ++a)
      for (int c = a; c < 3; ++c) amax = fmax(amax, fabs(A[a][c]));
    if (amax > 0.0 && amax < __builtin_huge_val()) {
      sc = scalbn(1.0, -ilogb(amax));
      for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) A[a][c] = A[a][c] * sc;
    }
#pragma unroll 1
    for (int sweep = 0; sweep < 6; ++sweep) {
      jacobi_rotate(A, V, 0, 1, 2, 3);
      jacobi_rotate(A, V, 0, 2, 1, 3);
      jacobi_rotate(A, V, 1, 2, 0, 3);
    }
  }
#pragma unroll
  for (int a = 0; a < D; ++a) mu[a] = A[a][a] / sc;
  return finite;
}

template <class T, int D>
__global__ void __launch_bounds__(kThreads) k_eigen_vjp(Geom g, const T* R1, const T* R2, const T* __restrict__ gn,
                                                        const T* __restrict__ gc, T* gR1, T* gR2,
                                                        const int32_t* __restrict__ on) {
  constexpr int K = D * (D + 1) / 2;
  const int64_t b = blockIdx.y;
  if (on != nullptr && on[b] == 0) return;
  const int64_t F = g.batch * g.count;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kThreads) + threadIdx.x; i < g.count;
       i += static_cast<int64_t>(gridDim.x) * kThreads) {
    const int64_t at = b * g.count + i;
    T r1[D], r2[K], tt[K];
#pragma unroll
    for (int a = 0; a < D; ++a) r1[a] = R1[a * F + at];
#pragma unroll
    for (int p = 0; p < K; ++p) r2[p] = R2[p * F + at];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int c = a; c < D; ++c) {
        T v = r1[a] * r1[c];
#pragma unroll
        for (int q = 0; q < D; ++q) v = v + r2[tensor_pair(D, a, q)] * r2[tensor_pair(D, q, c)];
        tt[tensor_pair(D, a, c)] = v;
      }
    double td[K], mu[3], V[3][3];
#pragma unroll
    for (int p = 0; p < K; ++p) td[p] = static_cast<double>(tt[p]);
    double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (eigen_pairs<D>(td, mu, V)) {
      int top = 0;
#pragma unroll
      for (int a = 1; a < D; ++a) top = mu[a] > mu[top] ? a : top;
      int sec = top == 0 ? 1 : 0;
#pragma unroll
      for (int a = 0; a < D; ++a) sec = (a != top && mu[a] > mu[sec]) ? a : sec;
      const double mu1 = mu[top], mu2 = mu[sec];
      double gmu1 = 0.0, gmu2 = 0.0;
      const double ssum = mu1 + mu2;
      const double gcv = static_cast<double>(gc[at]);
      if (ssum > 0.0) {
        const double craw = (mu1 - mu2) / ssum;
        if (craw >= 0.0 && craw <= 1.0) {
          gmu1 = gcv * (2.0 * mu2) / (ssum * ssum);
          gmu2 = -gcv * (2.0 * mu1) / (ssum * ssum);
        }
      }
      double gnv[3] = {0, 0, 0};
#pragma unroll
      for (int a = 0; a < D; ++a) gnv[a] = static_cast<double>(gn[a * F + at]);
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int c = 0; c < D; ++c) G[a][c] = gmu1 * V[a][top] * V[c][top] + gmu2 * V[a][sec] * V[c][sec];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        if (j == top) continue;
        const double gap = mu1 - mu[j];
        if (!(gap != 0.0)) continue;
        double dot = 0.0;
#pragma unroll
        for (int a = 0; a < D; ++a) dot += V[a][j] * gnv[a];
        const double coef = dot / gap;
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
          for (int c = 0; c < D; ++c) G[a][c] += 0.5 * coef * (V[a][j] * V[c][top] + V[a][top] * V[c][j]);
      }
    }
    // the pair gradients (an off-diagonal pair stands for two entries), then T's products
    double gp[K];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int c = a; c < D; ++c) gp[tensor_pair(D, a, c)] = a == c ? G[a][a] : G[a][c] + G[c][a];
    double g1[D], g2[K];
#pragma unroll
    for (int a = 0; a < D; ++a) g1[a] = 0.0;
#pragma unroll
    for (int p = 0; p < K; ++p) g2[p] = 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int c = a; c < D; ++c) {
        const double gv = gp[tensor_pair(D, a, c)];
        g1[a] += gv * static_cast<double>(r1[c]);
        g1[c] += gv * static_cast<double>(r1[a]);
#pragma unroll
        for (int q = 0; q < D; ++q) {
          g2[tensor_pair(D, a, q)] += gv * static_cast<double>(r2[tensor_pair(D, q, c)]);
          g2[tensor_pair(D, q, c)] += gv * static_cast<double>(r2[tensor_pair(D, a, q)]);
        }
      }
#pragma unroll
    for (int a = 0; a < D; ++a) gR1[a * F + at] = static_cast<T>(g1[a]);
#pragma unroll
    for (int p = 0; p < K; ++p) gR2[p * F + at] = static_cast<T>(g2[p]);
  }
}

// One component transposed: the forward took Y = T(m Xb) and R = Inv_odd(Y) (the inverse DCT with
// the sin basis on the odd axes). Inv_odd^T on an odd axis is the sum with sin(pi k (2x + 1) / 2n),
// which equals (-1)^x cos(pi (n - k) (2x + 1) / 2n): flip the signs of gR along the odd axes, take
// the forward DCT, read it at n_a - k_a on the odd axes (0 at k_a = 0), and gXb += m gY.
template <class T>
__global__ void __launch_bounds__(kThreads) k_riesz_acc(Geom g, const T* __restrict__ Z, T* __restrict__ gXb,
                                                        int kind, int a, int bb, int odd,
                                                        const int32_t* __restrict__ on) {
  const int64_t b = blockIdx.y;
  if (on != nullptr && on[b] == 0) return;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kThreads) + threadIdx.x; i < g.count;
       i += static_cast<int64_t>(gridDim.x) * kThreads) {
    int32_t k[kMaxDim];
    double u[kMaxDim];
    const double rad = dct_freq(g, i, k, u);
    const double m = riesz_mult(kind, a, bb, u, rad);
    int64_t src = 0;
    bool zero = false;
    for (int ax = 0; ax < g.dim; ++ax) {
      int32_t kk = k[ax];
      if ((odd >> ax) & 1) {
        if (kk == 0) zero = true;
        kk = g.n[ax] - kk;
      }
      src += static_cast<int64_t>(kk) * g.st[ax];
    }
    const int64_t at = b * g.count + i;
    const double z = zero ? 0.0 : static_cast<double>(Z[b * g.count + src]);
    gXb[at] = static_cast<T>(static_cast<double>(gXb[at]) + m * z);
  }
}

// gr += the gradient of r through the orientation, from gn (D fields) and gc; work: D + K + 4 fields.
template <class T>
void enqueue_orientation_vjp(hipStream_t s, int device, const Geom& g, const T* r, const int32_t* w,
                             const int32_t* on, const T* gn, const T* gc, T* gr, T* work, double* dscratch) {
  const int D = g.dim;
  if (D < 2) return;
  const int64_t F = g.batch * g.count;
  const int K = D * (D + 1) / 2;
  T* R1 = work;
  T* R2 = work + D * F;
  T* W3 = work + (D + K) * F;  // 3 fields for the forward's transforms, then gXb, Y, Z, scratch
  riesz_fields<T>(s, device, g, r, w, on, R1, R2, W3, dscratch);
  // the gradients of R1, R2 over R1, R2 themselves (each voxel reads its own first)
  const dim3 grid = dim3(static_cast<unsigned>(std::min<int64_t>((g.count + kThreads - 1) / kThreads, 1024)),
                         static_cast<unsigned>(g.batch));
  if (D == 2) EDA_GO((k_eigen_vjp<T, 2>), grid, dim3(kThreads), 0, s, g, R1, R2, gn, gc, R1, R2, on);
  else EDA_GO((k_eigen_vjp<T, 3>), grid, dim3(kThreads), 0, s, g, R1, R2, gn, gc, R1, R2, on);
  EDA_LAUNCH("k_eigen_vjp");
  T* gXb = W3;
  T* Y = W3 + F;
  T* Z = W3 + 2 * F;
  T* T1 = W3 + 3 * F;
  each(s, g, on, [=] __device__(int64_t b, int64_t i) { gXb[b * g.count + i] = T(0); });
  for (int ci = 0; ci < D + K; ++ci) {
    int a, bb, kind;
    riesz_component_of(D, ci, &a, &bb, &kind);
    const int odd = riesz_odd(kind, a, bb);
    const T* src = kind == 0 ? R1 + a * F : R2 + tensor_pair(D, a, bb) * F;
    each(s, g, on, [=] __device__(int64_t b, int64_t i) {
      int par = 0;
      for (int ax = 0; ax < g.dim; ++ax)
        if ((odd >> ax) & 1) par += static_cast<int>((i / g.st[ax]) % g.n[ax]);
      const int64_t at = b * g.count + i;
      Y[at] = (par & 1) ? -src[at] : src[at];
    });
    enqueue_dct<T>(s, device, g, Y, Z, T1, 0, 0);
    EDA_GO(k_riesz_acc<T>, grid, dim3(kThreads), 0, s, g, Z, gXb, kind, a, bb, odd, on);
    EDA_LAUNCH("k_riesz_acc");
  }
  // the band, then the transposed forward DCT (its inverse) and the mean's projection
  const double two_sigma2 = band_two_sigma2();
  each(s, g, on, [=] __device__(int64_t b, int64_t i) {
    int32_t k[kMaxDim];
    double u[kMaxDim];
    const double rad = dct_freq(g, i, k, u);
    double G = 0.0;
    if (rad > 0.0) {
      const double f0 = 1.0 / (1.0 * static_cast<double>(w[b]));
      const double l = log(rad / f0);
      G = exp(-(l * l) / two_sigma2);
    }
    const int64_t at = b * g.count + i;
    gXb[at] = static_cast<T>(static_cast<double>(gXb[at]) * G);
  });
  enqueue_dct<T>(s, device, g, gXb, Y, T1, 1, 0);
  double* mean = dscratch;
  enqueue_mean<T>(s, g, Y, on, mean, mean + g.batch);
  each(s, g, on, [=] __device__(int64_t b, int64_t i) {
    const int64_t at = b * g.count + i;
    gr[at] = static_cast<T>(static_cast<double>(gr[at]) + (static_cast<double>(Y[at]) - mean[b]));
  });
}

// The planes of the monogenic signal per voxel (D a template constant: the small arrays stay in
// registers): amplitude, phase, frequency, then (D >= 2) the direction with the canonical sign and
// the coherence.
template <class T, int D>
void monogenic_planes(hipStream_t s, const Geom& g, const T* f, const int32_t* kexp, double tiny, const T* R1,
                      const T* DF, const T* R2, const T* NV, const T* C, T* out, int64_t out_batch, int64_t b0) {
  const int64_t F = g.batch * g.count;
  const T zero = std::is_same_v<T, float> ? T(0x1p-14) : T(0x1p-30);
  const int64_t plane = out_batch * g.count;
  const T pi_t = static_cast<T>(3.141592653589793);
  each(s, g, nullptr, [=] __device__(int64_t b, int64_t i) {
    const int64_t at = b * g.count + i;
    T nt[D];
    if constexpr (D == 1) {
      nt[0] = T(1);
    } else {
      int sgn = 0;
#pragma unroll
      for (int a = 0; a < D; ++a) {
        nt[a] = NV[a * F + at];
        if (sgn == 0 && fabs(nt[a]) > zero) sgn = nt[a] > T(0) ? 1 : -1;
      }
#pragma unroll
      for (int a = 0; a < D; ++a) nt[a] = sgn < 0 ? -nt[a] : nt[a];
    }
    double n[D];
#pragma unroll
    for (int a = 0; a < D; ++a) n[a] = static_cast<double>(nt[a]);
    const double fv = static_cast<double>(f[at]);
    double q = 0.0, a2 = fv * fv, dnf = 0.0, dnq = 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const double r = static_cast<double>(R1[a * F + at]);
      q = q + n[a] * r;
      a2 = a2 + r * r;
      dnf = dnf + n[a] * static_cast<double>(DF[a * F + at]);
    }
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int c = a; c < D; ++c) {
        const double m = a == c ? n[a] * n[a] : 2.0 * (n[a] * n[c]);
        dnq = dnq + m * static_cast<double>(R2[(D == 1 ? 0 : tensor_pair(D, a, c)) * F + at]);
      }
    double phase = 0.0, freq = 0.0;
    if (a2 > tiny) {
      phase = atan2(q, fv);
      freq = (fv * dnq - q * dnf) / a2;
    }
    const int64_t ot = (b0 + b) * g.count + i;
    out[ot] = static_cast<T>(ldexp(sqrt(a2), kexp[b]));
    const T pt = static_cast<T>(phase);
    out[plane + ot] = pt == -pi_t ? pi_t : pt;
    out[2 * plane + ot] = static_cast<T>(freq);
    if constexpr (D >= 2) {
#pragma unroll
      for (int a = 0; a < D; ++a) out[3 * plane + ((b0 + b) * D + a) * g.count + i] = nt[a];
      out[(3 + D) * plane + ot] = C[at];
    }
  });
}

// The monogenic signal of every element of f (the prescaled fields f 2^-kexp[b]) into the planes
// of out (reference orientation.cu:420-493,789-831, monogenic.cpp:61-158).
template <class T>
void enqueue_monogenic(hipStream_t s, int device, const Geom& g, const T* f, const int32_t* kexp, double tiny,
                       T* work, double* dscratch, T* out, int64_t out_batch, int64_t b0) {
  const int D = g.dim;
  const int64_t F = g.batch * g.count;
  const int K = D >= 2 ? D * (D + 1) / 2 : 1;
  T* X = work;
  T* Y = work + F;
  T* T1 = work + 2 * F;
  T* R1 = work + 3 * F;
  T* DF = R1 + D * F;
  T* R2 = DF + D * F;  // R2, then T, then D_b R_a
  T* NV = R2 + K * F;
  T* C = NV + D * F;
  double* mean = dscratch;
  enqueue_mean<T>(s, g, f, nullptr, mean, mean + g.batch);
  each(s, g, nullptr, [=] __device__(int64_t b, int64_t i) {
    const int64_t at = b * g.count + i;
    Y[at] = static_cast<T>(static_cast<double>(f[at]) - mean[b]);
  });
  enqueue_dct<T>(s, device, g, Y, X, T1, 0, 0);
  for (int a = 0; a < D; ++a) riesz_component<T>(s, device, g, X, 0, a, a, Y, T1, R1 + a * F, nullptr);
  for (int a = 0; a < D; ++a) riesz_component<T>(s, device, g, X, 3, a, a, Y, T1, DF + a * F, nullptr);
  if (D >= 2) {
    for (int a = 0; a < D; ++a)
      for (int bb = a; bb < D; ++bb)
        riesz_component<T>(s, device, g, X, a == bb ? 1 : 2, a, bb, Y, T1, R2 + tensor_pair(D, a, bb) * F, nullptr);
    tensor_eigen<T>(s, g, R1, R2, nullptr, NV, C, nullptr);
    for (int a = 0; a < D; ++a)
      for (int bb = a; bb < D; ++bb)
        riesz_component<T>(s, device, g, X, 4, a, bb, Y, T1, R2 + tensor_pair(D, a, bb) * F, nullptr);
  } else {
    riesz_component<T>(s, device, g, X, 4, 0, 0, Y, T1, R2, nullptr);
  }
  // per voxel, in double: q = n . R1, A^2 = f^2 + |R1|^2, D_n f, D_n q; the phase and the frequency
  // where A^2 > tiny (else 0); the amplitude sqrt(A^2) 2^kexp rounded once; a phase that rounds to
  // -pi in T is pi
  if (D == 1) monogenic_planes<T, 1>(s, g, f, kexp, tiny, R1, DF, R2, NV, C, out, out_batch, b0);
  else if (D == 2) monogenic_planes<T, 2>(s, g, f, kexp, tiny, R1, DF, R2, NV, C, out, out_batch, b0);
  else monogenic_planes<T, 3>(s, g, f, kexp, tiny, R1, DF, R2, NV, C, out, out_batch, b0);
}

template const AxisMats<float>& axis_mats<float>(int, int32_t);
template const AxisMats<double>& axis_mats<double>(int, int32_t);
template void enqueue_dct<float>(hipStream_t, int, const Geom&, const float*, float*, float*, int, int);
template void enqueue_dct<double>(hipStream_t, int, const Geom&, const double*, double*, double*, int, int);
template void enqueue_eigen<float>(hipStream_t, const Geom&, const float*, const int32_t*, float*, float*, int32_t*);
template void enqueue_eigen<double>(hipStream_t, const Geom&, const double*, const int32_t*, double*, double*, int32_t*);
template void enqueue_orientation_vjp<float>(hipStream_t, int, const Geom&, const float*, const int32_t*,
                                             const int32_t*, const float*, const float*, float*, float*, double*);
template void enqueue_orientation_vjp<double>(hipStream_t, int, const Geom&, const double*, const int32_t*,
                                              const int32_t*, const double*, const double*, double*, double*, double*);
template void enqueue_orientation<float>(hipStream_t, int, const Geom&, const float*, const int32_t*, const int32_t*,
                                         float*, float*, float*, int32_t*, double*);
template void enqueue_orientation<double>(hipStream_t, int, const Geom&, const double*, const int32_t*, const int32_t*,
                                          double*, double*, double*, int32_t*, double*);
template void enqueue_monogenic<float>(hipStream_t, int, const Geom&, const float*, const int32_t*, double, float*,
                                       double*, float*, int64_t, int64_t);
template void enqueue_monogenic<double>(hipStream_t, int, const Geom&, const double*, const int32_t*, double, double*,
                                        double*, double*, int64_t, int64_t);

}  // namespace eda
