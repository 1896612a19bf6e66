// SPDX-License-Identifier: Apache-2.0
// Significant line extrema (reference cpp/src/method/stats_fields.cu:1-14,72-331): along every line
// of every axis with n >= 3, the peaks of r and of -r as scipy.signal.find_peaks defines them (a
// strict rise, a run of equal values, a strict fall; never a line end; a plateau's peak at its
// sample (left + right) // 2), whose topographic prominence (scipy.signal.peak_prominences) is at
// least tau at the peak. A side of the peak v is deep when some sample u before the first strictly
// higher one (or a NaN) has fl64(v - u) >= tau; the line end is not deep. Sets bit a of flags and
// counts the peaks into ext[3 b + a] by integer atomics (order independent).
//
// gfx950 mapping: one kernel writes the min and max of every 32-sample block of every line (lines
// longer than 32); then one thread per sample, the sample index fastest across the lanes along the
// contiguous axis and the adjacent lines fastest along the strided axes (coalesced either way).
// Only the thread of a run start owns a peak; it walks its own block sample by sample and then
// whole blocks by their summaries, entering a block only when it holds the side's end, so a side
// costs at most 63 + n / 32 steps.
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace eda {
namespace {

constexpr int kBlk = 32;

struct FlagLines {
  int64_t count;         // samples per element
  int32_t n;             // line length
  int32_t nb;            // blocks per line
  int64_t step, inner, outer;
  int64_t entries;       // outer * nb * inner block summaries per element (mins; the maxima follow)
  int32_t axis;
};

// 32-bit unsigned quotient (indices inside one element stay below 2^31).
__device__ __forceinline__ uint32_t udiv(uint32_t a, uint32_t b) { return a / b; }

template <class T>
__global__ void __launch_bounds__(kThreads) k_flag_blocks(FlagLines L, int64_t batch, const T* __restrict__ r,
                                                          T* __restrict__ sum, const int32_t* __restrict__ on) {
  const int64_t b = blockIdx.y;
  const uint32_t e = blockIdx.x * kThreads + threadIdx.x;
  if (e >= static_cast<uint64_t>(L.entries)) return;
  if (on != nullptr && on[b] == 0) return;
  const uint32_t inner = static_cast<uint32_t>(L.inner);
  const uint32_t t = udiv(e, inner);
  const uint32_t j = e - t * inner;
  const uint32_t o = udiv(t, static_cast<uint32_t>(L.nb));
  const int32_t c = static_cast<int32_t>(t - o * static_cast<uint32_t>(L.nb));
  const T* x = r + b * L.count + static_cast<int64_t>(o) * L.n * L.inner + j;
  const int32_t i0 = c * kBlk;
  const int32_t i1 = min(i0 + kBlk, L.n);
  T mn = x[i0 * L.step];
  T mx = mn;
  for (int32_t i = i0 + 1; i < i1; ++i) {
    const T v = x[i * L.step];
    mn = (v < mn || v != v) ? v : mn;
    mx = (v > mx || v != v) ? v : mx;
  }
  T* dst = sum + b * L.count + e;
  dst[0] = mn;
  dst[L.entries] = mx;
}

// fl64(a - b) >= tau (a float converts to double exactly).
template <class T>
__device__ __forceinline__ bool deep(T a, T b, T tau) {
  return static_cast<double>(a) - static_cast<double>(b) >= static_cast<double>(tau);
}
// One side sample u of the peak v: 2 ends the side not deep (higher, or NaN), 1 deep, 0 goes on.
template <bool Max, class T>
__device__ __forceinline__ int side_step(T u, T v, T tau) {
  if (Max ? !(u <= v) : !(u >= v)) return 2;
  return (Max ? deep<T>(v, u, tau) : deep<T>(u, v, tau)) ? 1 : 0;
}

template <bool Max, int Dir, class T>
__device__ bool side_deep(const T* __restrict__ x, int64_t st, int32_t n, int32_t nb, int32_t p, T v, T tau,
                          const T* __restrict__ bmin, const T* __restrict__ bmax, int64_t sst) {
  const int32_t cp = p / kBlk;
  const int32_t lim = Dir < 0 ? cp * kBlk : min(cp * kBlk + kBlk - 1, n - 1);
  for (int32_t k = p + Dir; Dir < 0 ? k >= lim : k <= lim; k += Dir) {
    const int sres = side_step<Max>(x[k * st], v, tau);
    if (sres) return sres == 1;
  }
  for (int32_t c = cp + Dir; Dir < 0 ? c >= 0 : c < nb; c += Dir) {
    const T hi = Max ? bmax[c * sst] : bmin[c * sst];
    if (Max ? !(hi <= v) : !(hi >= v)) {
      const int32_t k0 = Dir < 0 ? c * kBlk + kBlk - 1 : c * kBlk;
      const int32_t k1 = Dir < 0 ? c * kBlk : min(c * kBlk + kBlk - 1, n - 1);
      for (int32_t k = k0; Dir < 0 ? k >= k1 : k <= k1; k += Dir) {
        const int sres = side_step<Max>(x[k * st], v, tau);
        if (sres) return sres == 1;
      }
      return false;
    }
    const T lo = Max ? bmin[c * sst] : bmax[c * sst];
    if (Max ? deep<T>(v, lo, tau) : deep<T>(lo, v, tau)) return true;
  }
  return false;
}

template <class T>
__device__ int32_t run_end(const T* __restrict__ x, int64_t st, int32_t n, int32_t nb, int32_t s, T cur,
                           const T* __restrict__ bmin, const T* __restrict__ bmax, int64_t sst) {
  const int32_t lim = min((s / kBlk) * kBlk + kBlk - 1, n - 1);
  int32_t k = s + 1;
  while (k <= lim && x[k * st] == cur) ++k;
  if (k <= lim || k >= n) return k - 1;
  for (int32_t c = k / kBlk; c < nb; ++c) {
    if (bmin[c * sst] == cur && bmax[c * sst] == cur) continue;
    k = c * kBlk;
    while (x[k * st] == cur) ++k;
    return k - 1;
  }
  return n - 1;
}

template <class T>
__global__ void __launch_bounds__(kThreads) k_flags(FlagLines L, int64_t batch, const T* __restrict__ r,
                                                    const T* __restrict__ tau, uint8_t* __restrict__ flags,
                                                    uint64_t* __restrict__ ext, const int32_t* __restrict__ on,
                                                    const T* __restrict__ sum) {
  const int64_t b = blockIdx.y;
  if (on != nullptr && on[b] == 0) return;
  const int32_t n = L.n;
  const uint32_t inner = static_cast<uint32_t>(L.inner);
  uint64_t count = 0;
  for (int64_t e64 = blockIdx.x * static_cast<int64_t>(kThreads) + threadIdx.x; e64 < L.count;
       e64 += static_cast<int64_t>(gridDim.x) * kThreads) {
    const uint32_t e = static_cast<uint32_t>(e64);
    const uint32_t t = udiv(e, inner);
    const uint32_t j = e - t * inner;
    const uint32_t o = udiv(t, static_cast<uint32_t>(n));
    const int32_t s = static_cast<int32_t>(t - o * static_cast<uint32_t>(n));
    if (s < 1 || s > n - 2) continue;
    const int64_t base = b * L.count + static_cast<int64_t>(o) * n * L.inner + j;
    const T* x = r + base;
    const int64_t st = L.step;
    const T cur = x[s * st];
    const T prev = x[(s - 1) * st];
    if (prev < cur || prev > cur) {
      const T* bmin = sum + b * L.count + static_cast<int64_t>(o) * L.nb * L.inner + j;
      const T* bmax = bmin + L.entries;
      const int64_t sst = L.inner;
      const int32_t en = run_end(x, st, n, L.nb, s, cur, bmin, bmax, sst);
      if (en <= n - 2) {
        const T nxt = x[(en + 1) * st];
        const bool is_max = prev < cur && nxt < cur;
        const bool is_min = prev > cur && nxt > cur;
        if (is_max || is_min) {
          const int32_t p = (s + en) / 2;
          const T tp = tau[base + p * st];
          const bool sig = is_max ? side_deep<true, -1>(x, st, n, L.nb, p, cur, tp, bmin, bmax, sst) &&
                                        side_deep<true, 1>(x, st, n, L.nb, p, cur, tp, bmin, bmax, sst)
                                  : side_deep<false, -1>(x, st, n, L.nb, p, cur, tp, bmin, bmax, sst) &&
                                        side_deep<false, 1>(x, st, n, L.nb, p, cur, tp, bmin, bmax, sst);
          if (sig) {
            flags[base + p * st] |= static_cast<uint8_t>(1u << L.axis);
            ++count;
          }
        }
      }
    }
  }
  // the count: a wave64 sum, a workgroup sum, one integer atomic per workgroup
  __shared__ uint64_t red[kThreads / 64];
  count = wave_sum_u64(count);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = count;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t c = 0;
    for (int w = 0; w < kThreads / 64; ++w) c += red[w];
    if (c) atomicAdd(reinterpret_cast<unsigned long long*>(&ext[3 * b + L.axis]), static_cast<unsigned long long>(c));
  }
}

}  // namespace

template <class T>
void enqueue_sig_flags(hipStream_t s, const Geom& g, const T* r, const T* tau, uint8_t* flags, uint64_t* ext,
                       const int32_t* on, T* scratch) {
  each(s, g, on, [=] __device__(int64_t b, int64_t i) { flags[b * g.count + i] = 0; });
  for (int a = 0; a < g.dim; ++a) {
    const int32_t n = g.n[a];
    if (n < 3) continue;
    FlagLines L{};
    L.count = g.count;
    L.n = n;
    L.nb = (n + kBlk - 1) / kBlk;
    L.step = g.st[a];
    L.inner = g.st[a];
    L.outer = g.count / (static_cast<int64_t>(n) * L.inner);
    L.entries = L.outer * L.nb * L.inner;
    L.axis = a;
    if (n > kBlk) {
      hipLaunchKernelGGL(k_flag_blocks<T>,
                         dim3(static_cast<unsigned>((L.entries + kThreads - 1) / kThreads), static_cast<unsigned>(g.batch)),
                         dim3(kThreads), 0, s, L, g.batch, r, scratch, on);
      EDA_LAUNCH("k_flag_blocks");
    }
    // (workgroups per element, elements): at most 128 workgroups, so 128 atomics, per element
    int64_t gx = (g.count + kThreads - 1) / kThreads;
    gx = std::min<int64_t>(gx, std::max<int64_t>(16, std::min<int64_t>(128, 16384 / std::max<int64_t>(g.batch, 1))));
    hipLaunchKernelGGL(k_flags<T>, dim3(static_cast<unsigned>(gx), static_cast<unsigned>(g.batch)), dim3(kThreads), 0,
                       s, L, g.batch, r, tau, flags, ext, on, scratch);
    EDA_LAUNCH("k_flags");
  }
}

template void enqueue_sig_flags<float>(hipStream_t, const Geom&, const float*, const float*, uint8_t*, uint64_t*,
                                       const int32_t*, float*);
template void enqueue_sig_flags<double>(hipStream_t, const Geom&, const double*, const double*, uint8_t*, uint64_t*,
                                        const int32_t*, double*);

}  // namespace eda
