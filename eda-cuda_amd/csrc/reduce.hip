// SPDX-License-Identifier: Apache-2.0
// Canonical per-element reductions (reference cpp/src/field/common.cuh:134-259, field_ops.cu:52-83;
// host twin cpp/tests/twin.hpp:136-158): the energy sum v^2 in double, summed per last-axis row
// over 32-sample chunks sequentially and then by a pairwise tree over the chunk sums (padded with
// zeros to a power of two), and per element by the same tree over the row sums; max |v| as the
// integer maximum of the bit patterns of |v| (exact); the count of non-finite samples.
//
// gfx950 mapping: one wave64 per row, lane l owning chunks l, l + 64, l + 128, l + 192 (rows of up
// to 8192 samples); the tree levels 1..32 run across lanes by shuffles and the levels 64, 128
// across a lane's slots, which is the pairwise tree over chunk positions k * 64 + l. The element
// pass runs one workgroup per element over leaves of 1024 rows (a tree in LDS) combined by a
// binary-counter stack, equal to the pairwise tree over all rows.
#include "common.hpp"
#include "kernels.hpp"

namespace eda {
namespace {

constexpr int kRowSlots = 4;  // 4 x 64 chunks of 32 = 8192 samples

template <class T>
__global__ void __launch_bounds__(kThreads) k_row_stats(const T* __restrict__ f, int64_t rows_total, int32_t nl,
                                                        int64_t rows_per_elem, const int32_t* __restrict__ on,
                                                        double* __restrict__ row_e, uint64_t* __restrict__ row_a) {
  const int lane = threadIdx.x & 63;
  const int64_t q = blockIdx.x * static_cast<int64_t>(kThreads / 64) + (threadIdx.x >> 6);
  if (q >= rows_total) return;  // uniform per wave
  if (on != nullptr && on[q / rows_per_elem] == 0) return;
  const T* row = f + q * nl;
  const int32_t nchunks = (nl + 31) >> 5;
  double s[kRowSlots];
  uint64_t amax = 0, bad = 0;
#pragma unroll
  for (int k = 0; k < kRowSlots; ++k) {
    const int32_t c = k * 64 + lane;
    double acc = 0.0;
    if (c < nchunks) {
      const int32_t e0 = c * 32;
      const int32_t e1 = min(e0 + 32, nl);
      for (int32_t e = e0; e < e1; ++e) {
        const double d = static_cast<double>(row[e]);
        acc = acc + d * d;
        const uint64_t ab = abs_bits(row[e]);
        amax = ab > amax ? ab : amax;
        bad += isfinite(d) ? 0 : 1;
      }
    }
    s[k] = acc;
  }
#pragma unroll
  for (int k = 0; k < kRowSlots; ++k) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const double o = __shfl_down(s[k], d, 64);
      if ((lane & (2 * d - 1)) == 0) s[k] = s[k] + o;
    }
  }
  s[0] = s[0] + s[1];
  s[2] = s[2] + s[3];
  s[0] = s[0] + s[2];
  amax = wave_max_u64(amax);
  bad = wave_sum_u64(bad);
  if (lane == 0) {
    row_e[q] = s[0];
    row_a[2 * q] = amax;
    row_a[2 * q + 1] = bad;
  }
}

// Rows longer than 8192 samples (up to kLongChunks chunks of 32): one workgroup per row, the chunk
// sums in LDS, then the pairwise tree padded to a power of two in place (the canonical order).
constexpr int kLongChunks = 8192;

template <class T>
__global__ void __launch_bounds__(kThreads) k_row_stats_long(const T* __restrict__ f, int32_t nl, int64_t rows_per_elem,
                                                             const int32_t* __restrict__ on, double* __restrict__ row_e,
                                                             uint64_t* __restrict__ row_a) {
  __shared__ double cs[kLongChunks];
  __shared__ uint64_t red[2][kThreads / 64];
  const int64_t q = blockIdx.x;
  if (on != nullptr && on[q / rows_per_elem] == 0) return;
  const T* row = f + q * nl;
  const int32_t nchunks = (nl + 31) >> 5;
  int32_t m = 1;
  while (m < nchunks) m <<= 1;
  uint64_t amax = 0, bad = 0;
  for (int32_t c = threadIdx.x; c < m; c += kThreads) {
    double acc = 0.0;
    if (c < nchunks) {
      const int32_t e0 = c * 32, e1 = min(e0 + 32, nl);
      for (int32_t e = e0; e < e1; ++e) {
        const double d = static_cast<double>(row[e]);
        acc = acc + d * d;
        const uint64_t ab = abs_bits(row[e]);
        amax = ab > amax ? ab : amax;
        bad += isfinite(d) ? 0 : 1;
      }
    }
    cs[c] = acc;
  }
  __syncthreads();
  for (int32_t d = 1; d < m; d <<= 1) {
    for (int32_t i = threadIdx.x * 2 * d; i + d < m; i += kThreads * 2 * d) cs[i] = cs[i] + cs[i + d];
    __syncthreads();
  }
  amax = wave_max_u64(amax);
  bad = wave_sum_u64(bad);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = amax;
    red[1][threadIdx.x >> 6] = bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t mx = 0, c = 0;
    for (int w = 0; w < kThreads / 64; ++w) {
      mx = red[0][w] > mx ? red[0][w] : mx;
      c += red[1][w];
    }
    row_e[q] = cs[0];
    row_a[2 * q] = mx;
    row_a[2 * q + 1] = c;
  }
}

constexpr int kLeaf = 1024;

__global__ void __launch_bounds__(kThreads) k_elem_stats(int64_t rows_per_elem, const int32_t* __restrict__ on,
                                                         const double* __restrict__ row_e,
                                                         const uint64_t* __restrict__ row_a, double* energy,
                                                         uint64_t* amax_bits, uint64_t* nonfinite) {
  __shared__ double leaf[kLeaf];
  __shared__ uint64_t red[2][kThreads / 64];
  const int64_t b = blockIdx.x;
  if (on != nullptr && on[b] == 0) return;
  const int tid = threadIdx.x;
  const double* re = row_e + b * rows_per_elem;
  const uint64_t* ra = row_a + 2 * b * rows_per_elem;
  double stack[40];
  int level[40];
  int top = 0;
  uint64_t amax = 0, bad = 0;
  for (int64_t start = 0; start < rows_per_elem; start += kLeaf) {
    for (int i = tid; i < kLeaf; i += kThreads) {
      double v = 0.0;
      if (start + i < rows_per_elem) {
        v = re[start + i];
        const uint64_t a = ra[2 * (start + i)];
        amax = a > amax ? a : amax;
        bad += ra[2 * (start + i) + 1];
      }
      leaf[i] = v;
    }
    __syncthreads();
    for (int d = 1; d < kLeaf; d <<= 1) {
      for (int i = tid * 2 * d; i < kLeaf; i += kThreads * 2 * d) leaf[i] = leaf[i] + leaf[i + d];
      __syncthreads();
    }
    if (tid == 0) {
      double v = leaf[0];
      int l = 0;
      while (top > 0 && level[top - 1] == l) {
        v = stack[top - 1] + v;
        --top;
        ++l;
      }
      stack[top] = v;
      level[top] = l;
      ++top;
    }
    __syncthreads();
  }
  amax = wave_max_u64(amax);
  bad = wave_sum_u64(bad);
  if ((tid & 63) == 0) {
    red[0][tid >> 6] = amax;
    red[1][tid >> 6] = bad;
  }
  __syncthreads();
  if (tid == 0) {
    double v = 0.0;
    if (top > 0) {
      v = stack[top - 1];
      for (int i = top - 2; i >= 0; --i) v = stack[i] + v;
    }
    uint64_t m = 0, c = 0;
    for (int w = 0; w < kThreads / 64; ++w) {
      m = red[0][w] > m ? red[0][w] : m;
      c += red[1][w];
    }
    if (energy) energy[b] = v;
    if (amax_bits) amax_bits[b] = m;
    if (nonfinite) nonfinite[b] = c;
  }
}

}  // namespace

size_t elem_stats_ws_bytes(const Geom& g) {
  const int64_t rows = g.batch * (g.count / g.n[g.dim - 1]);
  return static_cast<size_t>(rows) * (sizeof(double) + 2 * sizeof(uint64_t)) + 256;
}

template <class T>
void enqueue_elem_stats(hipStream_t s, const Geom& g, const T* f, const int32_t* on, double* energy,
                        uint64_t* amax_bits, uint64_t* nonfinite, void* rows_ws) {
  const int32_t nl = g.n[g.dim - 1];
  const int64_t rpe = g.count / nl;
  const int64_t rows = g.batch * rpe;
  double* row_e = static_cast<double*>(rows_ws);
  uint64_t* row_a = reinterpret_cast<uint64_t*>(row_e + rows);
  if (nl > kRowSlots * 64 * 32) {
    if ((nl + 31) / 32 > kLongChunks) throw std::invalid_argument("eda: a row is too long for the reductions");
    EDA_GO(k_row_stats_long<T>, dim3(static_cast<unsigned>(rows)), dim3(kThreads), 0, s, f, nl, rpe, on, row_e,
           row_a);
    EDA_LAUNCH("k_row_stats_long");
  } else {
    const int64_t blocks = (rows + (kThreads / 64) - 1) / (kThreads / 64);
    EDA_GO(k_row_stats<T>, dim3(static_cast<unsigned>(blocks)), dim3(kThreads), 0, s, f, rows, nl, rpe, on,
                       row_e, row_a);
    EDA_LAUNCH("k_row_stats");
  }
  EDA_GO(k_elem_stats, dim3(static_cast<unsigned>(g.batch)), dim3(kThreads), 0, s, rpe, on, row_e, row_a,
                     energy, amax_bits, nonfinite);
  EDA_LAUNCH("k_elem_stats");
}

template void enqueue_elem_stats<float>(hipStream_t, const Geom&, const float*, const int32_t*, double*, uint64_t*,
                                        uint64_t*, void*);
template void enqueue_elem_stats<double>(hipStream_t, const Geom&, const double*, const int32_t*, double*, uint64_t*,
                                         uint64_t*, void*);

}  // namespace eda
