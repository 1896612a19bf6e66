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

// ---- the tiled pass: P lines of one axis staged in LDS -------------------------------------------------
// The side walk of side_deep on a line in LDS (stride 1), eight samples a loop iteration: each
// sample is tested branch-free in walk order (state 0 open, 1 deep, 2 ended not deep), so that a
// wave's lanes diverge once per eight samples, not per sample. Reads past the line's ends are
// clamped into it and their verdicts ignored.
template <bool Max, class T>
__device__ __forceinline__ int side_look(int st, T u, T v, double vd, double td) {
  const bool ends = Max ? !(u <= v) : !(u >= v);
  const bool dp = (Max ? vd - static_cast<double>(u) : static_cast<double>(u) - vd) >= td;
  return st != 0 ? st : (ends ? 2 : (dp ? 1 : 0));
}

// Samples first, first + Dir, ... (cnt of them) in walk order; returns the state.
template <bool Max, int Dir, class T>
__device__ __forceinline__ int walk8(const T* __restrict__ x, int32_t n, int32_t first, int32_t cnt, T v, double vd,
                                     double td) {
  int st = 0;
  for (int32_t j0 = 0; j0 < cnt; j0 += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int32_t k = min(max(first + Dir * (j0 + u), 0), n - 1);
      const int nst = side_look<Max>(st, x[k], v, vd, td);
      st = j0 + u < cnt ? nst : st;
    }
    if (st != 0) break;
  }
  return st;
}

template <bool Max, int Dir, class T>
__device__ bool side_deep_lds(const T* __restrict__ x, int32_t n, int32_t nb, int32_t p, T v, T tau,
                              const T* __restrict__ bmin, const T* __restrict__ bmax) {
  const double vd = static_cast<double>(v), td = static_cast<double>(tau);
  const int32_t cp = p / kBlk;
  const int32_t lim = Dir < 0 ? cp * kBlk : min(cp * kBlk + kBlk - 1, n - 1);
  const int st = walk8<Max, Dir>(x, n, p + Dir, Dir < 0 ? p - lim : lim - p, v, vd, td);
  if (st != 0) return st == 1;
  for (int32_t c = cp + Dir; Dir < 0 ? c >= 0 : c < nb; c += Dir) {
    const T hi = Max ? bmax[c] : bmin[c];
    if (Max ? !(hi <= v) : !(hi >= v)) {
      const int32_t k0 = Dir < 0 ? c * kBlk + kBlk - 1 : c * kBlk;
      const int32_t k1 = Dir < 0 ? c * kBlk : min(c * kBlk + kBlk - 1, n - 1);
      return walk8<Max, Dir>(x, n, k0, Dir < 0 ? k0 - k1 + 1 : k1 - k0 + 1, v, vd, td) == 1;
    }
    const T lo = Max ? bmin[c] : bmax[c];
    if (Max ? deep<T>(v, lo, tau) : deep<T>(lo, v, tau)) return true;
  }
  return false;
}

// A workgroup stages P lines (P consecutive rows of the contiguous axis, or P adjacent columns of a
// strided one) and their 32-sample block minima and maxima in LDS, finds and walks the peaks there
// with the same steps as k_flags (one thread per sample), marks them in an LDS byte tile, and writes
// the tile back over the flags (the first axis) or ORs it in (later axes), rows at a time.
struct FlagTile {
  int32_t n, nb, P, lpad;  // line length, blocks per line, lines per tile, LDS slots per line
  int64_t step, inner, outer, groups, count;
  int32_t axis, first, wide;  // first: the first axis pass (write, not OR); wide: 4-byte flag words
  int32_t xcd;                 // the XCD rotation's chunk (0: none)
};

template <class T>
__global__ void __launch_bounds__(kThreads) k_flags_tile(FlagTile a, const T* __restrict__ r,
                                                         const T* __restrict__ tau, uint8_t* __restrict__ flags,
                                                         uint64_t* __restrict__ ext, const int32_t* __restrict__ on) {
  extern __shared__ __align__(16) unsigned char smem[];
  // the XCD rotation: neighbouring column groups (which share cache lines) on one XCD
  const int64_t wg = a.xcd ? xcd_item(blockIdx.x, a.xcd) : static_cast<int64_t>(blockIdx.x);
  const int64_t b = wg / a.groups;
  const int64_t grp = wg - b * a.groups;
  if (on != nullptr && on[b] == 0) return;
  const int32_t n = a.n, P = a.P, nb = a.nb;
  T* X = reinterpret_cast<T*>(smem);  // line p at X + p lpad
  T* BMN = X + P * a.lpad;            // block c of line p at p nb + c
  T* BMX = BMN + P * nb;
  uint8_t* F = reinterpret_cast<uint8_t*>(BMX + P * nb);  // contiguous: p n + i; strided: i P + p
  const bool contiguous = a.inner == 1;
  int64_t base0;
  int32_t lines;
  if (contiguous) {
    const int64_t l0 = grp * P;
    lines = static_cast<int32_t>(min<int64_t>(P, a.outer - l0));
    base0 = b * a.count + l0 * n;
  } else {
    const int64_t gpo = (a.inner + P - 1) / P;
    const int64_t o = grp / gpo;
    const int64_t j0 = (grp - o * gpo) * P;
    lines = static_cast<int32_t>(min<int64_t>(P, a.inner - j0));
    base0 = b * a.count + o * static_cast<int64_t>(n) * a.inner + j0;
  }
  // sample i of line p in the field
  const auto gpos = [&](int32_t p, int32_t i) {
    return contiguous ? base0 + static_cast<int64_t>(p) * n + i : base0 + static_cast<int64_t>(i) * a.inner + p;
  };
  const auto fidx = [&](int32_t p, int32_t i) { return contiguous ? p * n + i : i * P + p; };
  const int32_t total = lines * n;
  const float inv_n = 1.0f / static_cast<float>(n);
  const int pshift = __ffs(P) - 1;  // strided: P is a power of two
  // stage: consecutive threads on consecutive samples of a row (coalesced either way)
  for (int32_t idx = threadIdx.x; idx < P * n; idx += kThreads) {
    int32_t p, i;
    if (contiguous) {
      p = static_cast<int32_t>(static_cast<float>(idx) * inv_n);
      i = idx - p * n;
      if (i < 0) { --p; i += n; } else if (i >= n) { ++p; i -= n; }
    } else {
      p = idx & (P - 1);
      i = idx >> pshift;
    }
    if (p < lines) X[p * a.lpad + i] = r[gpos(p, i)];
    F[fidx(p, i)] = 0;
  }
  __syncthreads();
  for (int32_t idx = threadIdx.x; idx < lines * nb; idx += kThreads) {
    const int32_t p = idx / nb;
    const int32_t c = idx - p * nb;
    const T* x = X + p * a.lpad;
    const int32_t i0 = c * kBlk;
    const int32_t i1 = min(i0 + kBlk, n);
    T mn = x[i0];
    T mx = mn;
    for (int32_t i = i0 + 1; i < i1; ++i) {
      const T v = x[i];
      mn = (v < mn || v != v) ? v : mn;
      mx = (v > mx || v != v) ? v : mx;
    }
    BMN[p * nb + c] = mn;
    BMX[p * nb + c] = mx;
  }
  __syncthreads();
  // the peaks of the tile into a list (one 16-bit entry each: pk << 4 | p << 1 | is_max), packed
  // per wave by a ballot, so that the side walks run on consecutive lanes
  uint16_t* PK = reinterpret_cast<uint16_t*>(F + ((P * n + 3) & ~3));
  __shared__ uint32_t npk;
  if (threadIdx.x == 0) npk = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  for (int32_t i0 = static_cast<int32_t>(threadIdx.x) - lane; i0 < total; i0 += kThreads) {
    const int32_t idx = i0 + lane;
    uint32_t entry = 0;
    bool peak = false;
    if (idx < total) {
      int32_t p = static_cast<int32_t>(static_cast<float>(idx) * inv_n);
      int32_t s = idx - p * n;
      if (s < 0) { --p; s += n; } else if (s >= n) { ++p; s -= n; }
      if (s >= 1 && s <= n - 2) {
        const T* x = X + p * a.lpad;
        const T cur = x[s];
        const T prev = x[s - 1];
        if (prev < cur || prev > cur) {
          const int32_t en = run_end(x, 1, n, nb, s, cur, BMN + p * nb, BMX + p * nb, 1);
          if (en <= n - 2) {
            const T nxt = x[en + 1];
            const bool is_max = prev < cur && nxt < cur;
            const bool is_min = prev > cur && nxt > cur;
            peak = is_max || is_min;
            entry = (static_cast<uint32_t>((s + en) / 2) << 4) | (static_cast<uint32_t>(p) << 1) | (is_max ? 1u : 0u);
          }
        }
      }
    }
    const uint64_t m = __ballot(peak);
    if (m != 0) {
      uint32_t at = 0;
      if (lane == 0) at = atomicAdd(&npk, static_cast<uint32_t>(__popcll(m)));
      at = static_cast<uint32_t>(__shfl(static_cast<int>(at), 0));
      if (peak) PK[at + static_cast<uint32_t>(__popcll(m & ((1ULL << lane) - 1ULL)))] = static_cast<uint16_t>(entry);
    }
  }
  __syncthreads();
  uint64_t count = 0;
  const uint8_t bit = static_cast<uint8_t>(1u << a.axis);
  const uint32_t np = npk;
  for (uint32_t q = threadIdx.x; q < np; q += kThreads) {
    const uint32_t e = PK[q];
    const int32_t pk = static_cast<int32_t>(e >> 4);
    const int32_t p = static_cast<int32_t>((e >> 1) & 7u);
    const T* x = X + p * a.lpad;
    const T* bmin = BMN + p * nb;
    const T* bmax = BMX + p * nb;
    const T cur = x[pk];  // a plateau's samples are equal
    const T tp = tau[gpos(p, pk)];
    const bool sig = (e & 1u) ? side_deep_lds<true, -1>(x, n, nb, pk, cur, tp, bmin, bmax) &&
                                    side_deep_lds<true, 1>(x, n, nb, pk, cur, tp, bmin, bmax)
                              : side_deep_lds<false, -1>(x, n, nb, pk, cur, tp, bmin, bmax) &&
                                    side_deep_lds<false, 1>(x, n, nb, pk, cur, tp, bmin, bmax);
    if (sig) {
      F[fidx(p, pk)] = bit;
      ++count;
    }
  }
  __syncthreads();
  // write-out: rows of the tile (a row: P adjacent columns, or a line), 4 flags a word where aligned
  if (a.wide) {
    const int32_t words = contiguous ? lines * n / 4 : n * P / 4;
    for (int32_t w = threadIdx.x; w < words; w += kThreads) {
      int64_t gp;
      if (contiguous) {
        gp = base0 + 4 * static_cast<int64_t>(w);  // lines are consecutive rows
      } else {
        const int32_t i = (4 * w) >> pshift;
        const int32_t p = 4 * w - (i << pshift);
        if (p >= lines) continue;
        gp = gpos(p, i);
      }
      const uint32_t f = *reinterpret_cast<const uint32_t*>(F + 4 * w);
      uint32_t* dst = reinterpret_cast<uint32_t*>(flags + gp);
      *dst = a.first ? f : (*dst | f);
    }
  } else {
    for (int32_t idx = threadIdx.x; idx < P * n; idx += kThreads) {
      int32_t p, i;
      if (contiguous) {
        p = idx / n;
        i = idx - p * n;
      } else {
        p = idx & (P - 1);
        i = idx >> pshift;
      }
      if (p >= lines) continue;
      const uint8_t f = F[fidx(p, i)];
      uint8_t* dst = flags + gpos(p, i);
      *dst = a.first ? f : static_cast<uint8_t>(*dst | f);
    }
  }
  __shared__ uint64_t red[kThreads / 64];
  count = wave_sum_u64(count);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = count;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t c = 0;
    for (int w = 0; w < kThreads / 64; ++w) c += red[w];
    if (c) atomicAdd(reinterpret_cast<unsigned long long*>(&ext[3 * b + a.axis]), static_cast<unsigned long long>(c));
  }
}

}  // namespace

template <class T>
void enqueue_sig_flags(hipStream_t s, const Geom& g, const T* r, const T* tau, uint8_t* flags, uint64_t* ext,
                       const int32_t* on, T* scratch) {
  // the tiled pass for lines up to 4096 samples: P lines a tile (a power of two on a strided axis)
  bool first = true;
  bool all_tiled = true;
  for (int a = 0; a < g.dim; ++a) all_tiled = all_tiled && (g.n[a] < 3 || g.n[a] <= 4096);
  if (all_tiled) {
    for (int a = 0; a < g.dim; ++a) {
      const int32_t n = g.n[a];
      if (n < 3) continue;
      FlagTile t{};
      t.n = n;
      t.nb = (n + kBlk - 1) / kBlk;
      t.step = g.st[a];
      t.inner = a == g.dim - 1 ? 1 : g.st[a];
      t.outer = g.count / (static_cast<int64_t>(n) * t.inner);
      t.count = g.count;
      t.axis = a;
      t.first = first ? 1 : 0;
      t.lpad = n | 1;
      int32_t P = 8;
      if (t.inner > 1) {
        while (P > 1 && P / 2 >= t.inner) P /= 2;
      } else {
        P = static_cast<int32_t>(std::min<int64_t>(P, t.outer));
      }
      // lines, block minima and maxima, the flag tile (rounded to 4 bytes), the peak list
      const auto bytes = [&](int32_t q) {
        return (static_cast<size_t>(q) * t.lpad + 2 * static_cast<size_t>(q) * t.nb) * sizeof(T) +
               (static_cast<size_t>(q) * n + 3) / 4 * 4 + 2 * static_cast<size_t>(q) * n;
      };
      while (P > 1 && bytes(P) > 40000) P /= 2;
      t.P = P;
      t.groups = t.inner == 1 ? (t.outer + P - 1) / P : t.outer * ((t.inner + P - 1) / P);
      const bool al4 = (reinterpret_cast<uintptr_t>(flags) & 3) == 0 && g.count % 4 == 0;
      t.wide = al4 && (t.inner == 1 ? n % 4 == 0 : (P % 4 == 0 && t.inner % 4 == 0)) ? 1 : 0;
      const int64_t wgs = g.batch * t.groups;
      if (wgs > 0x7fffffff) throw std::invalid_argument("eda: too many workgroups for the extrema pass");
      t.xcd = t.inner > 1 ? xcd_chunk(wgs) : 0;
      const bool timed = timer_family(kTimeFlags);
      hipEvent_t t0 = nullptr;
      if (timed) timer_note(s, &t0);
      EDA_GO(k_flags_tile<T>, dim3(static_cast<unsigned>(wgs)), dim3(kThreads), bytes(P), s, t, r, tau,
                         flags, ext, on);
      EDA_LAUNCH("k_flags_tile");
      if (timed) {
        // r and tau read, the flags read (after the first axis) and written, over the running elements
        int64_t running = on != nullptr ? mirror_active(on, g.batch) : g.batch;
        if (running < 0) running = g.batch;
        const double per = 2.0 * sizeof(T) + (first ? 1.0 : 2.0);
        timer_done(s, t0, per * static_cast<double>(running) * static_cast<double>(g.count), kTagFlags);
      }
      first = false;
    }
    if (!first) return;
  }
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
      EDA_GO(k_flag_blocks<T>,
                         dim3(static_cast<unsigned>((L.entries + kThreads - 1) / kThreads), static_cast<unsigned>(g.batch)),
                         dim3(kThreads), 0, s, L, g.batch, r, scratch, on);
      EDA_LAUNCH("k_flag_blocks");
    }
    // (workgroups per element, elements): at most 128 workgroups, so 128 atomics, per element
    int64_t gx = (g.count + kThreads - 1) / kThreads;
    gx = std::min<int64_t>(gx, std::max<int64_t>(16, std::min<int64_t>(128, 16384 / std::max<int64_t>(g.batch, 1))));
    EDA_GO(k_flags<T>, dim3(static_cast<unsigned>(gx), static_cast<unsigned>(g.batch)), dim3(kThreads), 0,
                       s, L, g.batch, r, tau, flags, ext, on, scratch);
    EDA_LAUNCH("k_flags");
  }
}

template void enqueue_sig_flags<float>(hipStream_t, const Geom&, const float*, const float*, uint8_t*, uint64_t*,
                                       const int32_t*, float*);
template void enqueue_sig_flags<double>(hipStream_t, const Geom&, const double*, const double*, uint8_t*, uint64_t*,
                                        const int32_t*, double*);

}  // namespace eda
