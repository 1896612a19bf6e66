// SPDX-License-Identifier: Apache-2.0
// Exact order statistics of per-element value sets, and the noise levels (reference
// cpp/src/method/stats.cu:1-347, stats_common.cuh:44-171).
//
// A selection finds the order statistics of ranks r0 <= r1 of every element's value set by six
// most-significant-digit radix passes (11, 11, 11, 11, 11, 9 bits) over the 64-bit order keys of
// the values as doubles (-0 counts as +0). Per pass a histogram kernel counts the digit of every
// value whose higher digits equal a rank's known prefix (LDS uint32 bins, flushed into the global
// 64-bit bins by integer atomics, so the counts never depend on the order of the atomics), and a
// pick kernel (one workgroup per element) finds the bin holding the rank. The median is the mean
// of the two middle values (numpy.median); the quantile follows numpy.quantile(method='linear'):
// virtual index (n - 1) q, the lerp that is exact at its end. An empty set gives NaN, count 0.
// The third pass also keeps the keys that match a rank's 22-bit prefix (the only values the later
// passes count) in a per-element candidate list; the last three passes count the list instead of
// reading every value again, unless it overflowed its slots.
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <cmath>

namespace eda {
namespace {

constexpr int kBits = 11;
constexpr int kBins = 1 << kBits;
constexpr int kPasses = 6;
constexpr int64_t kCandMax = int64_t{1} << 16;       // candidate slots per element at most
constexpr int64_t kCandTotal = int64_t{1} << 25;     // candidate slots of all elements at most (256 MiB)
constexpr uint32_t kBlockCand = 1024;                // candidates one workgroup gathers in LDS

// The candidate lists: element b's keys at keys + b slots, their number at n[b] (it counts past
// slots when the list overflows).
struct SelCand {
  uint64_t* keys;
  uint32_t* n;
  int64_t slots;
};

// Per element, the search state of the two ranks lo <= hi (median: the middle pair; quantile: the
// lerp's ends): key[r] holds the digits of rank r's value found so far, left[r] its rank among the
// values that share those digits. `merged` while both ranks share their digits (the histogram kernel
// then counts one row for both).
struct alignas(16) RankState {
  uint64_t key[2];
  uint64_t left[2];
  int64_t n;
  double frac;
  int32_t live;
  int32_t merged;
};

__host__ __device__ __forceinline__ uint64_t order_key(double d) {
  uint64_t u = dbits(d);
  if ((u << 1) == 0) u = 0;  // -0 -> +0
  return (u >> 63) ? ~u : (u | 0x8000000000000000ULL);
}
__device__ __forceinline__ double key_value(uint64_t k) {
  const uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffULL) : ~k;
  return bits_d(u);
}

template <class T>
struct SelArgs {
  Selection<T> sel;
  Geom g;
  int32_t naxes;
  int32_t ax[kMaxDim];
  int64_t items;
};

// The values of sample v of element b (every one goes to visit(double)): the field value, or the
// forward difference (in T) along each selected axis where v is not the axis's last sample;
// optionally |value - center[b]|. Indices inside one element stay below 2^31.
template <class T, class Visit>
__device__ __forceinline__ void sel_values(const SelArgs<T>& a, int64_t b, uint32_t v, Visit visit) {
  const int64_t i = b * a.g.count + v;
  if (a.sel.src == SelSrc::kField) {
    if (a.sel.mask != nullptr && (a.sel.mask[i] & a.sel.bits) == 0) return;
    double val = static_cast<double>(a.sel.field[i]);
    if (a.sel.center != nullptr) val = fabs(val - a.sel.center[b]);
    visit(val);
    return;
  }
  for (int k = 0; k < a.naxes; ++k) {
    const int axis = a.ax[k];
    const uint32_t st = static_cast<uint32_t>(a.g.st[axis]);
    const uint32_t nn = static_cast<uint32_t>(a.g.n[axis]);
    const uint32_t coord = axis == 0 ? v / st : (st == 1 ? v % nn : (v / st) % nn);
    if (coord >= nn - 1) continue;
    const T d = a.sel.field[i + st] - a.sel.field[i];  // the difference in T
    double val = static_cast<double>(d);
    if (a.sel.center != nullptr) val = fabs(val - a.sel.center[b]);
    visit(val);
  }
}

__global__ void k_sel_init(int64_t B, RankState* st, uint64_t* hist, uint32_t* ncand, const int32_t* on) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= B * 2 * kBins) return;
  const int64_t b = i / (2 * kBins);
  if (on != nullptr && on[b] == 0) return;
  hist[i] = 0;
  if (i % (2 * kBins) == 0) {
    st[b] = RankState{};
    ncand[b] = 0;
  }
}

// phase 0 counts every value; 1 also keeps the values that match a prefix as candidates; 2 counts
// the candidates when the list holds them all (else every value).
template <class T, int PHASE>
__global__ void __launch_bounds__(kThreads) k_sel_hist(SelArgs<T> a, const RankState* st, uint64_t* hist, int32_t nblk,
                                                       int32_t shift, int32_t width, int32_t first, const int32_t* on,
                                                       SelCand cand) {
  __shared__ uint32_t h[2][kBins];
  const int64_t b = blockIdx.x / nblk;
  const int64_t chunk = blockIdx.x % nblk;
  if (on != nullptr && on[b] == 0) return;
  const RankState S = st[b];
  if (!first && !S.live) return;
  const int nr = (first || S.merged) ? 1 : 2;
  for (int i = threadIdx.x; i < 2 * kBins; i += kThreads) h[i / kBins][i % kBins] = 0u;
  __syncthreads();
  const int hi = shift + width;
  const uint64_t dmask = (1ULL << width) - 1ULL;
  const uint32_t stride = static_cast<uint32_t>(nblk) * kThreads;
  const uint32_t count = static_cast<uint32_t>(a.g.count);
  // one count into bin of row r: where every active lane of the wave counts into the same bin (the
  // common case of the leading digits), the lowest one adds them all
  const int lane = threadIdx.x & 63;
  const auto count_in = [&](int r, uint32_t bin) {
    const uint64_t act = __ballot(1);
    const uint32_t b0 = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(bin)));
    if (__ballot(bin == b0) == act) {
      if (lane == __ffsll(static_cast<unsigned long long>(act)) - 1) atomicAdd(&h[r][b0], static_cast<uint32_t>(__popcll(act)));
    } else {
      atomicAdd(&h[r][bin], 1u);
    }
  };
  // counts key k where its higher digits match a rank's prefix; returns whether it matched one
  const auto tally = [&](uint64_t k) {
    const uint32_t bin = static_cast<uint32_t>((k >> shift) & dmask);
    if (first) {
      count_in(0, bin);
      return true;
    }
    const bool m0 = (k >> hi) == (S.key[0] >> hi);
    const bool m1 = nr == 2 && (k >> hi) == (S.key[1] >> hi);
    if (m0) count_in(0, bin);
    if (m1) count_in(1, bin);
    return m0 || m1;
  };
  const uint32_t nc = PHASE == 2 ? cand.n[b] : 0u;
  if (PHASE == 2 && nc <= static_cast<uint64_t>(cand.slots)) {
    const uint64_t* ck = cand.keys + b * cand.slots;
    for (uint32_t v = static_cast<uint32_t>(chunk) * kThreads + threadIdx.x; v < nc; v += stride) tally(ck[v]);
  } else if constexpr (PHASE == 1) {
    // the block's candidates gather in LDS (kBlockCand at most: more mark the list overflowed),
    // then one reservation in the element's list
    __shared__ uint64_t lk[kBlockCand];
    __shared__ uint32_t ln, lat;
    if (threadIdx.x == 0) ln = 0;
    __syncthreads();
    for (uint32_t v = static_cast<uint32_t>(chunk) * kThreads + threadIdx.x; v < count; v += stride)
      sel_values(a, b, v, [&](double val) {
        const uint64_t k = order_key(val);
        if (tally(k)) {
          const uint32_t at = atomicAdd(&ln, 1u);
          if (at < kBlockCand) lk[at] = k;
        }
      });
    __syncthreads();
    const uint32_t nl = ln;
    if (threadIdx.x == 0 && nl > 0)
      lat = atomicAdd(&cand.n[b], nl <= kBlockCand ? nl : static_cast<uint32_t>(cand.slots) + 1u);
    __syncthreads();
    if (nl <= kBlockCand && lat + nl <= static_cast<uint64_t>(cand.slots)) {
      uint64_t* ck = cand.keys + b * cand.slots + lat;
      for (uint32_t i = threadIdx.x; i < nl; i += kThreads) ck[i] = lk[i];
    }
  } else {
    for (uint32_t v = static_cast<uint32_t>(chunk) * kThreads + threadIdx.x; v < count; v += stride)
      sel_values(a, b, v, [&](double val) { tally(order_key(val)); });
  }
  __syncthreads();
  const int nb = 1 << width;
  for (int i = threadIdx.x; i < nr * nb; i += kThreads) {
    const int r = i / nb, bin = i % nb;
    const uint32_t c = h[r][bin];
    if (c) atomicAdd(reinterpret_cast<unsigned long long*>(&hist[(b * 2 + r) * kBins + bin]),
                     static_cast<unsigned long long>(c));
  }
}

// NumPy's linear quantile index: (n - 1) q; an index >= n - 1 takes the last value (gamma v + 1),
// an index < 0 the first.
__device__ __forceinline__ void quantile_index(int64_t n, double q, int64_t* i, int64_t* j, double* gamma) {
  const double v = static_cast<double>(n - 1) * q;
  if (v >= static_cast<double>(n - 1)) {
    *i = *j = n - 1;
    *gamma = v + 1.0;
  } else if (v < 0.0) {
    *i = *j = 0;
    *gamma = v;
  } else {
    const double f = floor(v);
    *i = static_cast<int64_t>(f);
    *j = *i + 1;
    *gamma = v - f;
  }
}

// The exclusive prefix sum of v over the workgroup's threads in thread order, and the total
// The inclusive prefix sum of v over the lanes of a wave (integers: exact in any order).
__device__ __forceinline__ uint64_t wave_inclusive(uint64_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t y = __shfl_up(v, d);
    v += lane >= d ? y : 0;
  }
  return v;
}

// One step of the rank search, one wave per element. Lane l owns the bins [l per, (l + 1) per) of
// the step's digit (per = 32 for 11-bit digits, 8 for the last 9 bits). The lanes' bin totals are
// scanned across the wave; the first lane whose running total passes a rank holds that rank's bin,
// walks its own bins to find it and broadcasts the bin and the rank left inside it. Every lane clears
// the bins it read (the next step counts into them again). After the last digit the wave writes the
// element's order statistic: the median as numpy.median (the mean of the middle pair), the quantile
// as numpy's 'linear' lerp; an empty value set gives NaN and count 0.
__global__ void __launch_bounds__(64) k_rank_step(RankState* st, uint64_t* hist, int32_t shift, int32_t width,
                                                  int32_t first, int32_t last, int32_t mode, double q,
                                                  const int32_t* on, double* out, int64_t* count) {
  constexpr int kPer = kBins / 64;  // bins per lane at most (32)
  const int64_t b = blockIdx.x;
  if (on != nullptr && on[b] == 0) return;
  const int lane = threadIdx.x;
  RankState S = st[b];
  if (!first && !S.live) return;
  const int per = (1 << width) >> 6;
  uint64_t* row[2] = {hist + (b * 2) * kBins, hist + (b * 2 + 1) * kBins};
  // the lane's bins of both rows, loaded once (all loads in flight together)
  const bool two = !first && !S.merged;
  uint64_t h[2][kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    h[0][j] = j < per ? row[0][lane * per + j] : 0;
    h[1][j] = two && j < per ? row[1][lane * per + j] : 0;
  }
  const auto lane_total = [&](int r) {
    uint64_t t = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) t += h[r][j];
    return t;
  };
  if (first) {
    const uint64_t n = static_cast<uint64_t>(__shfl(wave_inclusive(lane_total(0)), 63));
    S = RankState{};
    S.n = static_cast<int64_t>(n);
    S.live = n > 0 ? 1 : 0;
    S.merged = 1;
    if (n == 0) {
#pragma unroll
      for (int j = 0; j < kPer; ++j)
        if (j < per) row[0][lane * per + j] = 0;
      if (lane == 0) {
        st[b] = S;
        out[b] = __longlong_as_double(0x7ff8000000000000LL);
        if (count != nullptr) count[b] = 0;
      }
      return;
    }
    if (mode == 0) {
      S.left[0] = (n - 1) / 2;
      S.left[1] = n / 2;
    } else {
      int64_t i = 0, j = 0;
      quantile_index(static_cast<int64_t>(n), q, &i, &j, &S.frac);
      S.left[0] = static_cast<uint64_t>(i);
      S.left[1] = static_cast<uint64_t>(j);
    }
  }
  for (int r = 0; r < 2; ++r) {
    const int hr = S.merged ? 0 : r;
    const uint64_t tot = lane_total(hr);
    const uint64_t incl = wave_inclusive(tot);
    const uint64_t want = S.left[r];
    const uint64_t past = __ballot(want < incl);  // lanes whose running total passes the rank
    const int owner = __ffsll(static_cast<unsigned long long>(past)) - 1;
    uint32_t bin = 0;
    uint64_t rest = 0;
    if (lane == owner) {
      uint64_t below = incl - tot;  // the running total before the lane
      bool found = false;
#pragma unroll
      for (int j = 0; j < kPer; ++j) {
        const uint64_t c = h[hr][j];
        if (!found && j < per && want < below + c) {
          bin = static_cast<uint32_t>(lane * per + j);
          rest = want - below;
          found = true;
        }
        below += c;
      }
    }
    bin = static_cast<uint32_t>(__shfl(static_cast<int>(bin), owner));
    rest = static_cast<uint64_t>(__shfl(static_cast<unsigned long long>(rest), owner));
    S.key[r] |= static_cast<uint64_t>(bin) << shift;
    S.left[r] = rest;
  }
  // clear the bins this step read (the next step counts into them again)
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    if (j < per) {
      row[0][lane * per + j] = 0;
      row[1][lane * per + j] = 0;
    }
  }
  S.merged = S.key[0] == S.key[1] ? 1 : 0;
  if (lane != 0) return;
  st[b] = S;
  if (!last) return;
  if (count != nullptr) count[b] = S.n;
  const double lo = key_value(S.key[0]);
  const double hi = key_value(S.key[1]);
  if (mode == 0) {
    out[b] = (S.n & 1) ? lo : (lo + hi) / 2.0;
  } else if (S.frac >= 0.5) {
    out[b] = hi - (hi - lo) * (1.0 - S.frac);
  } else {
    out[b] = lo + (hi - lo) * S.frac;
  }
}

size_t align_up(size_t v) { return (v + 255) / 256 * 256; }


// Candidate slots per element (0: the passes read every value): the values of the largest
// selection (every axis's differences) up to kCandMax, within kCandTotal over the batch.
int64_t cand_slots(const Geom& g) {
  const int64_t items = static_cast<int64_t>(g.dim) * g.count;
  const int64_t c = std::min({items, kCandMax, kCandTotal / std::max<int64_t>(1, g.batch)});
  return c < 64 ? 0 : c;
}

// the selection's state, histograms, candidate counts and lists
size_t sel_state_bytes(const Geom& g) { return align_up(static_cast<size_t>(g.batch) * sizeof(RankState)); }
size_t sel_hist_bytes(const Geom& g) { return align_up(static_cast<size_t>(g.batch) * 2 * kBins * sizeof(uint64_t)); }
size_t sel_count_bytes(const Geom& g) { return align_up(static_cast<size_t>(g.batch) * sizeof(uint32_t)); }

}  // namespace

int64_t select_cand_slots(const Geom& g) { return cand_slots(g); }
size_t select_ws_bytes(const Geom& g, int64_t slots) {
  if (slots < 0) slots = cand_slots(g);
  return sel_state_bytes(g) + sel_hist_bytes(g) + sel_count_bytes(g) +
         align_up(static_cast<size_t>(g.batch) * static_cast<size_t>(slots) * sizeof(uint64_t));
}
size_t noise_ws_bytes(const Geom& g) { return select_ws_bytes(g) + 3 * align_up(static_cast<size_t>(g.batch) * 8); }

template <class T>
void enqueue_select(hipStream_t s, const Geom& g, const Selection<T>& sel, int mode, double q, double* out,
                    int64_t* count, const int32_t* on, void* ws, int64_t slots) {
  SelArgs<T> a{};
  a.sel = sel;
  a.g = g;
  if (sel.src == SelSrc::kField) {
    a.naxes = 1;
    a.items = g.count;
  } else {
    for (int ax = 0; ax < g.dim; ++ax)
      if (sel.axes & (1u << ax)) a.ax[a.naxes++] = ax;
    a.items = static_cast<int64_t>(a.naxes) * g.count;
  }
  const int64_t B = g.batch;
  char* w = static_cast<char*>(ws);
  RankState* st = reinterpret_cast<RankState*>(w);
  uint64_t* hist = reinterpret_cast<uint64_t*>(w + sel_state_bytes(g));
  SelCand cand{};
  cand.n = reinterpret_cast<uint32_t*>(w + sel_state_bytes(g) + sel_hist_bytes(g));
  cand.keys = reinterpret_cast<uint64_t*>(w + sel_state_bytes(g) + sel_hist_bytes(g) + sel_count_bytes(g));
  cand.slots = slots < 0 ? cand_slots(g) : slots;
  const int64_t ni = B * 2 * kBins;
  EDA_GO(k_sel_init, dim3(static_cast<unsigned>((ni + kThreads - 1) / kThreads)), dim3(kThreads), 0, s, B,
                     st, hist, cand.n, on);
  EDA_LAUNCH("k_sel_init");
  int64_t nblk = (g.count + int64_t{kThreads} * 32 - 1) / (int64_t{kThreads} * 32);  // blocks per element
  nblk = nblk < 1 ? 1 : (nblk > 512 ? 512 : nblk);
  if (a.items > 0) {
    for (int p = 0; p < kPasses; ++p) {
      const int top = 64 - kBits * p;
      const int shift = top - kBits > 0 ? top - kBits : 0;
      const int width = top - shift;
      const int phase = cand.slots == 0 ? 0 : p == 2 ? 1 : p > 2 ? 2 : 0;
      const dim3 grid(static_cast<unsigned>(nblk * B));
      const int32_t fst = p == 0 ? 1 : 0;
      if (phase == 1)
        EDA_GO((k_sel_hist<T, 1>), grid, dim3(kThreads), 0, s, a, st, hist, static_cast<int32_t>(nblk), shift,
               width, fst, on, cand);
      else if (phase == 2)
        EDA_GO((k_sel_hist<T, 2>), grid, dim3(kThreads), 0, s, a, st, hist, static_cast<int32_t>(nblk), shift,
               width, fst, on, cand);
      else
        EDA_GO((k_sel_hist<T, 0>), grid, dim3(kThreads), 0, s, a, st, hist, static_cast<int32_t>(nblk), shift,
               width, fst, on, cand);
      EDA_LAUNCH("k_sel_hist");
      EDA_GO(k_rank_step, dim3(static_cast<unsigned>(B)), dim3(64), 0, s, st, hist, shift, width, fst,
             p == kPasses - 1 ? 1 : 0, mode, q, on, out, count);
      EDA_LAUNCH("k_rank_step");
    }
  } else {
    flat(s, B, [=] __device__(int64_t b) {
      if (on != nullptr && on[b] == 0) return;
      out[b] = __longlong_as_double(0x7ff8000000000000LL);
      if (count != nullptr) count[b] = 0;
    });
  }
}

// sigma = (1.4826 MAD(d)) / sqrt 2 (reference stats.cu:262-334): pooled over the axes, or the
// largest over the axes (the first axis wins ties).
template <class T>
void enqueue_noise_sigma(hipStream_t s, const Geom& g, const T* r, bool per_axis, double* out, const int32_t* on,
                         void* ws) {
  const int64_t B = g.batch;
  char* extra = static_cast<char*>(ws) + select_ws_bytes(g);
  double* center = reinterpret_cast<double*>(extra);
  double* mad = reinterpret_cast<double*>(extra + align_up(static_cast<size_t>(B) * 8));
  double* acc = reinterpret_cast<double*>(extra + 2 * align_up(static_cast<size_t>(B) * 8));
  const double sqrt2 = std::sqrt(2.0);
  Selection<T> sel{};
  sel.src = SelSrc::kDiff;
  sel.field = r;
  const auto level = [&](uint8_t axes, int op) {
    sel.axes = axes;
    sel.center = nullptr;
    enqueue_select<T>(s, g, sel, 0, 0.5, center, nullptr, on, ws);
    sel.center = center;
    enqueue_select<T>(s, g, sel, 0, 0.5, mad, nullptr, on, ws);
    flat(s, B, [=] __device__(int64_t b) {
      if (on != nullptr && on[b] == 0) return;
      const double v = (1.4826 * mad[b]) / sqrt2;
      if (op == 0) out[b] = v;
      else if (v > acc[b]) acc[b] = v;
    });
  };
  if (!per_axis) {
    level(static_cast<uint8_t>((1u << g.dim) - 1u), 0);
    return;
  }
  flat(s, B, [=] __device__(int64_t b) { acc[b] = 0.0; });
  for (int a = 0; a < g.dim; ++a) {
    if (g.n[a] < 2) continue;
    level(static_cast<uint8_t>(1u << a), 1);
  }
  flat(s, B, [=] __device__(int64_t b) {
    if (on == nullptr || on[b] != 0) out[b] = acc[b];
  });
}

template void enqueue_select<float>(hipStream_t, const Geom&, const Selection<float>&, int, double, double*, int64_t*,
                                    const int32_t*, void*, int64_t);
template void enqueue_select<double>(hipStream_t, const Geom&, const Selection<double>&, int, double, double*, int64_t*,
                                     const int32_t*, void*, int64_t);
template void enqueue_noise_sigma<float>(hipStream_t, const Geom&, const float*, bool, double*, const int32_t*, void*);
template void enqueue_noise_sigma<double>(hipStream_t, const Geom&, const double*, bool, double*, const int32_t*,
                                          void*);

}  // namespace eda
