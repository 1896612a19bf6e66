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
// The reference computes the transforms with cuFFT through Makhoul's permutation; this build
// contracts every axis with its dense DCT matrix instead, as batched GEMMs: the DCT of an n-sample
// axis is an (n x n) matrix, so a transform of a (rows x n) block is one GEMM, with no library and
// no permutation passes. The matrices are built on the device from cos / sin tables of
// pi j / (2n), j < 4n, evaluated in long double on the host.
#include "common.hpp"
#include "kernels.hpp"

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

template <class T>
void gemm(hipStream_t s, const GemmDims& d, const T* A, const T* B, T* C, const int32_t* on) {
  const dim3 grid(static_cast<unsigned>((d.N + kTN - 1) / kTN), static_cast<unsigned>((d.M + kTM - 1) / kTM),
                  static_cast<unsigned>(d.batches));
  const bool timed = timer_gemm_on();
  hipEvent_t t0 = nullptr;
  if (timed) timer_note(s, &t0);
  if constexpr (std::is_same_v<T, float>) {
    hipLaunchKernelGGL(k_gemm_mfma, grid, dim3(kThreads), 0, s, d, A, B, C, on);
    EDA_LAUNCH("k_gemm_mfma");
  } else {
    hipLaunchKernelGGL(k_gemm<T>, grid, dim3(kThreads), 0, s, d, A, B, C, on);
    EDA_LAUNCH("k_gemm");
  }
  if (timed) timer_done(s, t0, 2.0 * static_cast<double>(d.M) * d.N * d.K * d.batches);
}

// One axis of a separable transform: out = M applied along axis a of every running element
// (M row-major (out index, in index), n x n).
template <class T>
void axis_transform(hipStream_t s, const Geom& g, int a, const T* M, const T* in, T* out, const int32_t* on) {
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
  flat(nullptr, static_cast<int64_t>(nn), [=] __device__(int64_t e) {
    const int64_t k = e / n, x = e - k * n;
    const int64_t q = (k * (2 * x + 1)) % N4;
    fwd[k * n + x] = static_cast<T>(dck[k] * dcs[q]);
    icos[x * n + k] = static_cast<T>(dck[k] * dcs[q]);
    isin[x * n + k] = static_cast<T>(dck[k] * dsn[q]);
  });
  EDA_HIP(hipDeviceSynchronize());
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
  hipLaunchKernelGGL(k_mean_part<T>, dim3(kMeanParts, static_cast<unsigned>(g.batch)), dim3(kThreads), 0, s, r, g.count,
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
  // the axes from last to first, ping-ponging between tmp and out so that the last pass lands in out
  const int D = g.dim;
  const T* src = in;
  for (int k = 0; k < D; ++k) {
    const int a = D - 1 - k;
    const AxisMats<T>& m = axis_mats<T>(device, g.n[a]);
    const T* M = !inverse ? m.fwd : (((odd >> a) & 1) ? m.isin : m.icos);
    T* dst = ((D - 1 - k) % 2 == 0) ? out : tmp;
    axis_transform<T>(s, g, a, M, src, dst, nullptr);
    src = dst;
  }
}

template <class T>
void enqueue_eigen(hipStream_t s, const Geom& g, const T* t, const int32_t* on, T* nvec, T* coh, int32_t* nonfinite) {
  if (g.dim == 2) eigen_pass<T, 2>(s, g, t, on, nvec, coh, nonfinite);
  else if (g.dim == 3) eigen_pass<T, 3>(s, g, t, on, nvec, coh, nonfinite);
}

// T = R1 R1^T + R2 R2^T per voxel, written over R2 (pair order), then the eigen pass into n and c.
template <class T>
void tensor_eigen(hipStream_t s, const Geom& g, const T* R1, T* R2, const int32_t* on, T* nvec, T* coh,
                  int32_t* nonfinite) {
  const int D = g.dim;
  const int K = D * (D + 1) / 2;
  const int64_t F = g.batch * g.count;
  each(s, g, on, [=] __device__(int64_t b, int64_t i) {
    const int64_t at = b * g.count + i;
    T r1[kMaxDim], r2[6], out[6];
    for (int a = 0; a < D; ++a) r1[a] = R1[a * F + at];
    for (int p = 0; p < K; ++p) r2[p] = R2[p * F + at];
    for (int a = 0; a < D; ++a)
      for (int c = a; c < D; ++c) {
        T v = r1[a] * r1[c];
        for (int q = 0; q < D; ++q) v = v + r2[tensor_pair(D, a, q)] * r2[tensor_pair(D, q, c)];
        out[tensor_pair(D, a, c)] = v;
      }
    for (int p = 0; p < K; ++p) R2[p * F + at] = out[p];
  });
  enqueue_eigen<T>(s, g, R2, on, nvec, coh, nonfinite);
}

// One component of X (a kind of riesz_mult) into dst: Y = T(m X), then the inverse transform with
// the sin basis on the component's odd axes (T1: the transform's scratch field).
template <class T>
void riesz_component(hipStream_t s, int device, const Geom& g, const T* X, int kind, int a, int bb, T* Y, T* T1,
                     T* dst, const int32_t* on) {
  each(s, g, on, [=] __device__(int64_t b, int64_t i) {
    int32_t k[kMaxDim];
    double u[kMaxDim];
    const double rad = dct_freq(g, i, k, u);
    const double m = riesz_mult(kind, a, bb, u, rad);
    const int64_t at = b * g.count + i;
    Y[at] = static_cast<T>(m * static_cast<double>(X[at]));
  });
  enqueue_dct<T>(s, device, g, Y, dst, T1, 1, riesz_odd(kind, a, bb));
}

template <class T>
void enqueue_orientation(hipStream_t s, int device, const Geom& g, const T* r, const int32_t* w, const int32_t* on,
                         T* nvec, T* coh, T* work, int32_t* nonfinite, double* dscratch) {
  const int D = g.dim;
  if (D < 2) return;
  const int64_t F = g.batch * g.count;
  const int K = D * (D + 1) / 2;
  T* X = work;
  T* Y = work + F;
  T* T1 = work + 2 * F;
  T* R2 = work + 3 * F;
  T* R1 = nvec;  // n is written last, over R1
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
    // ci < D: R1_ci; then the pairs (a, bb), a <= bb, in the pair order
    int a = ci, bb = ci;
    if (ci >= D) {
      int p = ci - D;
      for (a = 0; a < D; ++a) {
        if (p < D - a) break;
        p -= D - a;
      }
      bb = a + p;
    }
    const int kind = ci < D ? 0 : (a == bb ? 1 : 2);
    riesz_component<T>(s, device, g, X, kind, a, bb, Y, T1, kind == 0 ? R1 + a * F : R2 + tensor_pair(D, a, bb) * F,
                       on);
  }
  tensor_eigen<T>(s, g, R1, R2, on, nvec, coh, nonfinite);
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
  const T zero = std::is_same_v<T, float> ? T(0x1p-14) : T(0x1p-30);
  const int64_t plane = out_batch * g.count;
  const T pi_t = static_cast<T>(3.141592653589793);
  each(s, g, nullptr, [=] __device__(int64_t b, int64_t i) {
    const int64_t at = b * g.count + i;
    T nt[kMaxDim];
    if (D == 1) {
      nt[0] = T(1);
    } else {
      int sgn = 0;
      for (int a = 0; a < D; ++a) {
        nt[a] = NV[a * F + at];
        if (sgn == 0 && fabs(nt[a]) > zero) sgn = nt[a] > T(0) ? 1 : -1;
      }
      for (int a = 0; a < D; ++a) nt[a] = sgn < 0 ? -nt[a] : nt[a];
    }
    double n[kMaxDim];
    for (int a = 0; a < D; ++a) n[a] = static_cast<double>(nt[a]);
    const double fv = static_cast<double>(f[at]);
    double q = 0.0, a2 = fv * fv, dnf = 0.0, dnq = 0.0;
    for (int a = 0; a < D; ++a) {
      const double r = static_cast<double>(R1[a * F + at]);
      q = q + n[a] * r;
      a2 = a2 + r * r;
      dnf = dnf + n[a] * static_cast<double>(DF[a * F + at]);
    }
    int p = 0;
    for (int a = 0; a < D; ++a)
      for (int c = a; c < D; ++c, ++p) {
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
    if (D >= 2) {
      for (int a = 0; a < D; ++a) out[3 * plane + ((b0 + b) * D + a) * g.count + i] = nt[a];
      out[(3 + D) * plane + ot] = C[at];
    }
  });
}

template const AxisMats<float>& axis_mats<float>(int, int32_t);
template const AxisMats<double>& axis_mats<double>(int, int32_t);
template void enqueue_dct<float>(hipStream_t, int, const Geom&, const float*, float*, float*, int, int);
template void enqueue_dct<double>(hipStream_t, int, const Geom&, const double*, double*, double*, int, int);
template void enqueue_eigen<float>(hipStream_t, const Geom&, const float*, const int32_t*, float*, float*, int32_t*);
template void enqueue_eigen<double>(hipStream_t, const Geom&, const double*, const int32_t*, double*, double*, int32_t*);
template void enqueue_orientation<float>(hipStream_t, int, const Geom&, const float*, const int32_t*, const int32_t*,
                                         float*, float*, float*, int32_t*, double*);
template void enqueue_orientation<double>(hipStream_t, int, const Geom&, const double*, const int32_t*, const int32_t*,
                                          double*, double*, double*, int32_t*, double*);
template void enqueue_monogenic<float>(hipStream_t, int, const Geom&, const float*, const int32_t*, double, float*,
                                       double*, float*, int64_t, int64_t);
template void enqueue_monogenic<double>(hipStream_t, int, const Geom&, const double*, const int32_t*, double, double*,
                                        double*, double*, int64_t, int64_t);

}  // namespace eda
