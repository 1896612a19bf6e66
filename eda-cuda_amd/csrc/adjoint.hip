// SPDX-License-Identifier: Apache-2.0
// The adjoint of the decomposition (eda_backward) and its frozen forward (eda_replay).
//
// With the decisions of a call frozen (its tape, tape.hpp), the rows are a piecewise-smooth
// function of the input: per mode h = S(r) and r' = r - h, rows = h 2^k and the residue. The
// backward runs the modes in reverse; for each it recomputes the mode's forward fields from the
// taped residual with the forward's own kernels (nothing is stored across modes) and pulls the
// gradient back through every stage:
//   - the element operators: the rung cubes are recomputed; the gather's transpose scatters each
//     voxel's gradient onto the rung fields (max / min: to the one of the 5 points that won; box:
//     to all 5, over the D-linear corners of the interpolated ones) and collects the gradient of
//     the hat weight t and of the segment vector o; the cube's transpose runs the axis passes
//     backwards: a max / min pass sends each output's gradient to the position that attains the
//     window's extreme (the leftmost one where values tie), a box pass is its own mirror with the
//     half-sample reflection folded back;
//   - G1's cubes, ramps, the floor, the quotient, the blend and the sift update pointwise;
//   - the orientation: t and o depend on the coherence c and direction n (d = L ((1 - c) + c / q),
//     e = 0.5 max(L (1 + (span - 1) c) - d, 0), o = e n, t = hat(log d)); n and c on the structure
//     tensor's eigen decomposition (first-order perturbation), the tensor on the Riesz components,
//     and those on r through the transposed inverse transforms and the transposed DCT (orient.hip).
// The forward stages follow reference cpp/src/method/driver.cpp:540-689, element_op.cuh:86-282,
// sift.cu:55-232, vhgw.cu:281-452 and orientation.cu:86-404; the reference has no adjoint.
#include "common.hpp"
#include "kernels.hpp"
#include "tape.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

namespace eda {
namespace {

// ---- the transposed axis passes ------------------------------------------------------------------
// A tile of P lines along `axis` of one element: adjacent rows (contiguous axis) or adjacent columns
// (strided axis), staged in LDS as [p][i].
struct LineT {
  Geom g;
  int32_t axis, n, P, rungs;
  int64_t inner, outer, tiles, F;
  int64_t vstride;     // rung stride of vals (0: one input for every rung)
  int32_t sum_rungs;   // 1: the rungs' gradients summed into gx (+= when add); 0: rung j into gx + j F
  int32_t add;
  int32_t scale;       // box: times 1 / prod over the axes of the clamped sides
  int32_t lds_stride;  // LDS words per staged line (odd)
  int32_t scr_stride;  // transposed max / min: uint16 slots of a wave's H and G positions
  int32_t pk_stride;   // transposed max / min: uint16 slots of a wave's picks
  int32_t xcd;         // workgroups in XCD-rotated chunks of this many tiles (0: in order)
};

__device__ __forceinline__ int64_t line_base(const LineT& a, int64_t b, int64_t tile, int32_t p, bool* ok) {
  if (a.inner == 1) {
    const int64_t line = tile * a.P + p;
    *ok = line < a.outer;
    return b * a.g.count + line * a.n;
  }
  const int64_t cb = (a.inner + a.P - 1) / a.P;
  const int64_t o = tile / cb;
  const int64_t c = (tile - o * cb) * a.P + p;
  *ok = c < a.inner;
  return b * a.g.count + o * static_cast<int64_t>(a.n) * a.inner + c;
}

// The bases of the tile's P lines (<= 16), computed once per workgroup (-1: a line outside the
// tile); the staging loops then use 32-bit index math only.
constexpr int kLineTileMax = 16;
__device__ __forceinline__ void tile_bases(const LineT& a, int64_t b, int64_t tile, int64_t* sbase) {
  if (threadIdx.x < a.P) {
    bool ok;
    const int64_t base = line_base(a, b, tile, static_cast<int32_t>(threadIdx.x), &ok);
    sbase[threadIdx.x] = ok ? base : -1;
  }
  __syncthreads();
}

// loads rows of a rung field into s[p][i] (zeros for lines outside the tile); 8 reads per thread in
// flight at a time (clamped indices: every read is of a valid address, its value then selected)
__device__ __forceinline__ int32_t slot32(int32_t i) { return i + (i >> 5); }

template <class T, bool SLOTS = false>
__device__ __forceinline__ void stage(const LineT& a, const int64_t* sbase, const T* __restrict__ src, T* s, T mul) {
  const int32_t total = a.P * a.n;
  const int64_t step = a.inner == 1 ? 1 : a.inner;
  const int32_t nthr = blockDim.x;
  const auto where = [&](int32_t k, int32_t* p, int32_t* i) {
    if (a.inner == 1) {
      *p = k / a.n;
      *i = k - *p * a.n;
    } else {
      *i = k / a.P;
      *p = k - *i * a.P;
    }
  };
  for (int32_t k0 = threadIdx.x; k0 < total; k0 += 8 * nthr) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      int32_t p, i;
      where(min(k0 + u * nthr, total - 1), &p, &i);
      const int64_t base = sbase[p];
      v[u] = src[(base >= 0 ? base : 0) + i * (base >= 0 ? step : 0)];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int32_t k = k0 + u * nthr;
      if (k < total) {
        int32_t p, i;
        where(k, &p, &i);
        s[p * a.lds_stride + (SLOTS ? slot32(i) : i)] = sbase[p] >= 0 ? v[u] * mul : T(0);
      }
    }
  }
}

template <class T, bool SLOTS = false>
__device__ __forceinline__ void unstage(const LineT& a, const int64_t* sbase, T* __restrict__ dst, const T* s,
                                        bool add) {
  const int32_t total = a.P * a.n;
  const int64_t step = a.inner == 1 ? 1 : a.inner;
  for (int32_t k = threadIdx.x; k < total; k += blockDim.x) {
    int32_t p, i;
    if (a.inner == 1) {
      p = k / a.n;
      i = k - p * a.n;
    } else {
      i = k / a.P;
      p = k - i * a.P;
    }
    const int64_t base = sbase[p];
    if (base < 0) continue;
    const int64_t at = base + i * step;
    const T v = s[p * a.lds_stride + (SLOTS ? slot32(i) : i)];
    dst[at] = add ? dst[at] + v : v;
  }
}

// The transposed max (MIN: min) pass of rungs j0 .. j1 - 1. Each output x sends its gradient to the
// position of its window's extreme, the leftmost where values tie (the window clipped to the line,
// which holds the same values as the half-sample reflected one; -0 and +0 tie).
//
// One wave per line, the line's values, gradient and sums staged in LDS (slot i + i / 32, so that
// lanes eight samples apart fall in distinct banks). The extreme positions come from van Herk /
// Gil-Werman folds over blocks of L = 2r + 1 samples aligned at 0, carried as (order key, position)
// pairs in registers: each lane folds a run of 8 consecutive samples, the folds cross lanes by a
// segmented scan (DPP row shifts, readlane across rows), and the suffix folds H and prefix folds G
// keep only their positions. Output x picks the better of H(max(0, x - r)) and G(min(n - 1, x + r)).
// The picks do not decrease with x, so the outputs that pick one position form one run of
// consecutive x: a segmented sum over the runs (in double; lane runs, then a scan across lanes)
// leaves each run's gradient at its last output, which adds it at the picked position.
template <class T>
struct ExtKey {
  okey_t<T> k;
  int32_t i;
};
constexpr int kDppRowShl = 0x100, kDppRowShr = 0x110, kDppWaveShl1 = 0x130, kDppWaveShr1 = 0x138;
template <int CTRL>
__device__ __forceinline__ int32_t dpp_i32(int32_t old, int32_t v) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, 0xf, 0xf, false);
}
template <int CTRL, class S>
__device__ __forceinline__ S dpp_any(S old, S v) {
  if constexpr (sizeof(S) == 4) {
    return __builtin_bit_cast(S, dpp_i32<CTRL>(__builtin_bit_cast(int32_t, old), __builtin_bit_cast(int32_t, v)));
  } else {
    const uint64_t o = __builtin_bit_cast(uint64_t, old), x = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = static_cast<uint32_t>(dpp_i32<CTRL>(static_cast<int32_t>(o), static_cast<int32_t>(x)));
    const uint32_t hi = static_cast<uint32_t>(dpp_i32<CTRL>(static_cast<int32_t>(o >> 32), static_cast<int32_t>(x >> 32)));
    return __builtin_bit_cast(S, (static_cast<uint64_t>(hi) << 32) | lo);
  }
}
template <class S>
__device__ __forceinline__ S readlane_any(S v, int l) {
  if constexpr (sizeof(S) == 4) {
    return __builtin_bit_cast(S, __builtin_amdgcn_readlane(__builtin_bit_cast(int32_t, v), l));
  } else {
    const uint64_t x = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int32_t>(x), l));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int32_t>(x >> 32), l));
    return __builtin_bit_cast(S, (static_cast<uint64_t>(hi) << 32) | lo);
  }
}
template <int CTRL, class T>
__device__ __forceinline__ ExtKey<T> dpp_key(ExtKey<T> old, ExtKey<T> v) {
  return {dpp_any<CTRL>(old.k, v.k), dpp_i32<CTRL>(old.i, v.i)};
}
template <class T>
__device__ __forceinline__ ExtKey<T> readlane_key(ExtKey<T> v, int l) {
  return {readlane_any(v.k, l), __builtin_amdgcn_readlane(v.i, l)};
}
// The better of two pairs: the larger (MIN: smaller) key, the smaller position where keys tie.
template <class T, bool MIN>
__device__ __forceinline__ ExtKey<T> key_best(ExtKey<T> a, ExtKey<T> b) {
  const bool ta = MIN ? (a.k < b.k || (a.k == b.k && a.i < b.i)) : (a.k > b.k || (a.k == b.k && a.i < b.i));
  return ta ? a : b;
}
template <class T, bool MIN>
__device__ __forceinline__ ExtKey<T> key_ident() {
  using K = okey_t<T>;
  return {MIN ? std::numeric_limits<K>::max() : std::numeric_limits<K>::lowest(), 0x7fffffff};
}
// the order key of v with -0 mapped to +0
template <class T>
__device__ __forceinline__ okey_t<T> ext_okey(T v) {
  return okey(v == T(0) ? T(0) : v);
}
__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// Inclusive segmented scan across the wave: a lane's (a, f) becomes the fold of a over the lanes
// from the nearest lane with f set to itself, in lane order (UP) or reverse lane order; depth: the
// doubling steps inside a 16-lane row, then the rows in order.
template <bool UP, class V, class Op>
__device__ __forceinline__ void lane_scan(V& a, int& f, int depth, V id, Op op) {
  const auto step = [&](auto dc) {
    constexpr int ctrl = UP ? kDppRowShr + decltype(dc)::value : kDppRowShl + decltype(dc)::value;
    V a2;
    if constexpr (std::is_same_v<V, ExtKey<float>> || std::is_same_v<V, ExtKey<double>>) a2 = dpp_key<ctrl>(id, a);
    else a2 = dpp_any<ctrl>(id, a);
    const int f2 = dpp_i32<ctrl>(0, f);
    if (!f) {
      a = UP ? op(a2, a) : op(a, a2);
      f = f2;
    }
  };
  if (depth > 0) step(std::integral_constant<int, 1>{});
  if (depth > 1) step(std::integral_constant<int, 2>{});
  if (depth > 2) step(std::integral_constant<int, 4>{});
  if (depth > 3) step(std::integral_constant<int, 8>{});
  const int row = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int k = 1; k < 4; ++k) {
    const int rr = UP ? k : 3 - k;
    const int src = UP ? 16 * rr - 1 : 16 * rr + 16;
    V as;
    if constexpr (std::is_same_v<V, ExtKey<float>> || std::is_same_v<V, ExtKey<double>>) as = readlane_key(a, src);
    else as = readlane_any(a, src);
    const int fs = __builtin_amdgcn_readlane(f, src);
    if (row == rr && !f) {
      a = UP ? op(as, a) : op(a, as);
      f = fs;
    }
  }
}

constexpr int kExtKT = 8;                  // samples per lane
constexpr int32_t kExtTile = 64 * kExtKT;  // samples per wave pass

// One line: V the values, Gy the output gradient (slots slot32(i)); Acc (same slots) += the
// line's input gradient; Hi, Gi (slots slot32(i)) and Pk (plain) the wave's scratch.
template <class T, bool MIN>
__device__ void ext_line_vjp(const T* __restrict__ V, const T* __restrict__ Gy, T* __restrict__ Acc,
                             uint16_t* __restrict__ Hi, uint16_t* __restrict__ Gi, uint16_t* __restrict__ Pk,
                             int32_t n, int32_t r) {
  constexpr int KT = kExtKT;
  using E = ExtKey<T>;
  const int lane = threadIdx.x & 63;
  const int32_t L = 2 * r + 1;
  const float inv = 1.0f / static_cast<float>(L);
  uint32_t pat = 0;
  for (int32_t j = 0; j < KT; j += L) pat |= 1u << j;
  const int32_t chain = (L - 1) / KT;
  const int depth = chain == 0 ? 0 : chain < 2 ? 1 : chain < 4 ? 2 : chain < 8 ? 3 : 4;
  const int32_t ntiles = (n + kExtTile - 1) / kExtTile;
  const E id = key_ident<T, MIN>();
  const auto best = [](E a, E b) { return key_best<T, MIN>(a, b); };
  const auto starts_of = [&](int32_t i0) {
    int32_t m0 = i0 - static_cast<int32_t>(static_cast<float>(i0) * inv) * L;
    m0 += m0 < 0 ? L : (m0 >= L ? -L : 0);
    const int32_t k0 = m0 == 0 ? 0 : L - m0;
    return k0 < KT ? (pat << k0) & ((1u << KT) - 1u) : 0u;
  };
  // G: prefix folds inside the blocks, left to right
  E carry = id;
  for (int32_t t = 0; t < ntiles; ++t) {
    const int32_t i0 = t * kExtTile + lane * KT;
    const uint32_t starts = starts_of(i0);
    E gv[KT];
    E g = id;
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const int32_t x = i0 + k;
      const E e = x < n ? E{ext_okey(V[slot32(x)]), x} : id;
      g = (starts >> k) & 1u ? e : best(g, e);
      gv[k] = g;
    }
    E a = gv[KT - 1];
    int f = starts != 0;
    if (chain > 0) lane_scan<true>(a, f, depth, id, best);
    if (!f) a = best(carry, a);
    const E prv = dpp_key<kDppWaveShr1>(carry, a);
    const int kf = starts ? __ffs(static_cast<int>(starts)) - 1 : KT;
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const E v = k < kf ? best(prv, gv[k]) : gv[k];
      if (i0 + k < n) Gi[slot32(i0 + k)] = static_cast<uint16_t>(v.i);
    }
    carry = readlane_key(a, 63);
  }
  // H: suffix folds inside the blocks, right to left
  carry = id;
  for (int32_t t = ntiles - 1; t >= 0; --t) {
    const int32_t i0 = t * kExtTile + lane * KT;
    const uint32_t starts = starts_of(i0);
    int32_t me = i0 + KT - static_cast<int32_t>(static_cast<float>(i0 + KT) * inv) * L;
    me += me < 0 ? L : (me >= L ? -L : 0);
    const uint32_t ends = (starts >> 1) | (me == 0 ? 1u << (KT - 1) : 0u);
    E hv[KT];
    E h = id;
#pragma unroll
    for (int k = KT - 1; k >= 0; --k) {
      const int32_t x = i0 + k;
      const E e = x < n ? E{ext_okey(V[slot32(x)]), x} : id;
      h = (ends >> k) & 1u ? e : best(e, h);
      hv[k] = h;
    }
    E a = hv[0];
    int f = ends != 0;
    if (chain > 0) lane_scan<false>(a, f, depth, id, best);
    if (!f) a = best(a, carry);
    const E nxt = dpp_key<kDppWaveShl1>(carry, a);
    const int kl = ends ? 31 - __clz(static_cast<int>(ends)) : -1;
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const E v = k > kl ? best(hv[k], nxt) : hv[k];
      if (i0 + k < n) Hi[slot32(i0 + k)] = static_cast<uint16_t>(v.i);
    }
    carry = readlane_key(a, 0);
  }
  wave_fence();
  // the picks
  const int32_t mlast = (n - 1) - static_cast<int32_t>(static_cast<float>(n - 1) * inv) * L;
  const int32_t last = (n - 1) - (mlast < 0 ? mlast + L : (mlast >= L ? mlast - L : mlast));  // the last block's start
  for (int32_t x = lane; x < n; x += 64) {
    const int32_t lo = x - r > 0 ? x - r : 0;
    const int32_t hi = x + r < n - 1 ? x + r : n - 1;
    int32_t p;
    if (x < r) {
      p = Gi[slot32(hi)];
    } else if (x + r > n - 1 && lo >= last) {
      p = Hi[slot32(lo)];
    } else {
      const int32_t ph = Hi[slot32(lo)], pg = Gi[slot32(hi)];
      const okey_t<T> kh = ext_okey(V[slot32(ph)]), kg = ext_okey(V[slot32(pg)]);
      p = (MIN ? kh <= kg : kh >= kg) ? ph : pg;  // ph < pg unless both are the whole block's extreme
    }
    Pk[x] = static_cast<uint16_t>(p);
  }
  wave_fence();
  // segmented sums of the gradient over the runs of equal picks, left to right
  double cs = 0.0;    // the open run's sum at the end of the previous tile
  int32_t cp = -1;    // the previous tile's last pick
  for (int32_t t = 0; t < ntiles; ++t) {
    const int32_t i0 = t * kExtTile + lane * KT;
    int32_t p[KT];
    double s[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) p[k] = i0 + k < n ? static_cast<int32_t>(Pk[i0 + k]) : 0x10000;
    const int32_t pl = dpp_i32<kDppWaveShr1>(cp, p[KT - 1]);
    uint32_t st = 0;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const bool s0 = p[k] != (k ? p[k - 1] : pl);
      st |= s0 ? 1u << k : 0u;
      const double gv = i0 + k < n ? static_cast<double>(Gy[slot32(i0 + k)]) : 0.0;
      acc = s0 ? gv : acc + gv;
      s[k] = acc;
    }
    double a = s[KT - 1];
    int f = st != 0;
    lane_scan<true>(a, f, 4, 0.0, [](double u, double v) { return u + v; });
    if (!f) a = cs + a;
    const double prv = dpp_any<kDppWaveShr1>(cs, a);
    const int kf = st ? __ffs(static_cast<int>(st)) - 1 : KT;
#pragma unroll
    for (int k = 0; k < KT; ++k)
      if (k < kf) s[k] = prv + s[k];
    // the last output of a run adds the run's sum at its pick
    const int32_t after = i0 + KT < n ? static_cast<int32_t>(Pk[i0 + KT]) : 0x10000;
    const int32_t pn = dpp_i32<kDppWaveShl1>(after, p[0]);
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const int32_t x = i0 + k;
      const int32_t q = k + 1 < KT ? p[k + 1] : pn;
      if (x < n && (x == n - 1 || q != p[k])) Acc[slot32(p[k])] += static_cast<T>(s[k]);
    }
    cs = readlane_any(a, 63);
    cp = __builtin_amdgcn_readlane(p[KT - 1], 63);
  }
  wave_fence();
}

// The workgroup stages P lines of the tile (values once when every rung reads the same input, else
// per rung; the output gradient per rung), and its waves take the lines in turn.
template <class T, bool MIN>
__global__ void __launch_bounds__(512) k_line_extT(LineT a, const T* __restrict__ vals, const T* __restrict__ gy,
                                                   T* __restrict__ gx, const int32_t* __restrict__ win,
                                                   const int32_t* __restrict__ on) {
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ int64_t sbase[kLineTileMax];
  const int32_t st = a.lds_stride;
  const int32_t S = a.P * st;
  T* sv = reinterpret_cast<T*>(lds);
  T* sg = sv + S;
  T* sacc = sg + S;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int W = blockDim.x >> 6;
  const int32_t sst = a.scr_stride;  // uint16 slots of Hi / Gi; Pk: ntiles * 512
  uint16_t* scr = reinterpret_cast<uint16_t*>(sacc + S) + static_cast<int64_t>(wave) * (2 * sst + a.pk_stride);
  uint16_t* Hi = scr;
  uint16_t* Gi = scr + sst;
  uint16_t* Pk = scr + 2 * sst;
  const int64_t wg = a.xcd ? xcd_item(blockIdx.x, a.xcd) : static_cast<int64_t>(blockIdx.x);
  const int64_t b = wg / a.tiles;
  const int64_t tile = wg - b * a.tiles;
  if (on != nullptr && on[b] == 0) return;
  tile_bases(a, b, tile, sbase);
  const int32_t j0 = a.sum_rungs ? 0 : static_cast<int32_t>(blockIdx.y);
  const int32_t j1 = a.sum_rungs ? a.rungs : j0 + 1;
  const int32_t n = a.n;
  (void)lane;
  for (int32_t k = threadIdx.x; k < S; k += blockDim.x) sacc[k] = T(0);
  bool have_v = false;
  for (int32_t j = j0; j < j1; ++j) {
    const int32_t w = win[static_cast<int64_t>(j) * a.g.batch + b];
    if (w < 1) continue;
    const int32_t r = min((w - 1) / 2, n - 1);
    __syncthreads();
    if (!have_v || a.vstride != 0) stage<T, true>(a, sbase, vals + j * a.vstride, sv, T(1));
    have_v = true;
    stage<T, true>(a, sbase, gy + j * a.F, sg, T(1));
    __syncthreads();
    for (int32_t p = wave; p < a.P; p += W)
      if (sbase[p] >= 0) ext_line_vjp<T, MIN>(sv + p * st, sg + p * st, sacc + p * st, Hi, Gi, Pk, n, r);
  }
  __syncthreads();
  if (a.sum_rungs) unstage<T, true>(a, sbase, gx, sacc, a.add != 0);
  else unstage<T, true>(a, sbase, gx + j0 * a.F, sacc, false);
}

// The transposed box pass: y(x) = sum over p in [x - r, x + r] of v(reflect(p)) gives
// gx(q) = sum over the three pre-images c of q under the reflection (q, -q - 1, 2n - 1 - q) of the
// window sums of gy (zero outside the line) centred at c, from prefix sums in double.
template <class T>
__global__ void __launch_bounds__(kThreads) k_line_boxT(LineT a, const T* __restrict__ gy, T* __restrict__ gx,
                                                        const int32_t* __restrict__ win,
                                                        const int32_t* __restrict__ on) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int32_t S = a.P * a.lds_stride;
  double* pre = reinterpret_cast<double*>(lds);  // P (n + 1) prefix sums, stride lds_stride + 1
  T* sg = reinterpret_cast<T*>(pre + a.P * (a.lds_stride + 1));
  T* sacc = sg + S;
  __shared__ int64_t sbase[kLineTileMax];
  const int64_t b = blockIdx.x / a.tiles;
  const int64_t tile = blockIdx.x - b * a.tiles;
  if (on != nullptr && on[b] == 0) return;
  tile_bases(a, b, tile, sbase);
  const int32_t j0 = a.sum_rungs ? 0 : static_cast<int32_t>(blockIdx.y);
  const int32_t j1 = a.sum_rungs ? a.rungs : j0 + 1;
  const int32_t n = a.n;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int32_t k = threadIdx.x; k < S; k += blockDim.x) sacc[k] = T(0);
  for (int32_t j = j0; j < j1; ++j) {
    const int32_t w = win[static_cast<int64_t>(j) * a.g.batch + b];
    if (w < 1) continue;
    const int32_t r = min((w - 1) / 2, n - 1);
    T fac = T(1);
    if (a.scale) {
      double prod = 1.0;
      for (int ax = 0; ax < a.g.dim; ++ax) prod *= static_cast<double>(2 * min((w - 1) / 2, a.g.n[ax] - 1) + 1);
      fac = static_cast<T>(1.0 / prod);
    }
    __syncthreads();
    stage<T>(a, sbase, gy + j * a.F, sg, fac);
    __syncthreads();
    // prefix sums, one wave per line: lane chunks, then a wave scan of the chunk totals
    const int32_t chunk = (n + 63) / 64;
    for (int32_t p = wave; p < a.P; p += blockDim.x / 64) {
      const T* gl = sg + p * a.lds_stride;
      double* pl = pre + p * (a.lds_stride + 1);
      const int32_t i0 = lane * chunk;
      double tot = 0.0;
      for (int32_t i = i0; i < min(i0 + chunk, n); ++i) tot += static_cast<double>(gl[i]);
      double incl = tot;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const double y = __shfl_up(incl, d, 64);
        if (lane >= d) incl += y;
      }
      double run = incl - tot;
      if (lane == 0) pl[0] = 0.0;
      for (int32_t i = i0; i < min(i0 + chunk, n); ++i) {
        run += static_cast<double>(gl[i]);
        pl[i + 1] = run;
      }
    }
    __syncthreads();
    for (int32_t k = threadIdx.x; k < a.P * n; k += blockDim.x) {
      const int32_t p = k / n, q = k - p * n;
      const double* pl = pre + p * (a.lds_stride + 1);
      double s = 0.0;
      const int32_t cs[3] = {q, -q - 1, 2 * n - 1 - q};
#pragma unroll
      for (int u = 0; u < 3; ++u) {
        const int32_t lo = max(cs[u] - r, 0), hi = min(cs[u] + r + 1, n);
        if (hi > lo) s += pl[hi] - pl[lo];
      }
      sacc[p * a.lds_stride + q] += static_cast<T>(s);
    }
  }
  __syncthreads();
  if (a.sum_rungs) unstage<T>(a, sbase, gx, sacc, a.add != 0);
  else unstage<T>(a, sbase, gx + j0 * a.F, sacc, false);
}

// The longest axis the transposed passes take (P n samples staged per workgroup).
constexpr int32_t kLineMax = 2048;

template <class T>
LineT line_shape(const Geom& g, int axis, int32_t rungs, int64_t F, bool box) {
  LineT a{};
  a.g = g;
  a.axis = axis;
  a.n = g.n[axis];
  if (a.n > kLineMax)
    throw std::invalid_argument("eda: the backward takes axes of at most " + std::to_string(kLineMax) +
                                " samples, got " + std::to_string(a.n));
  a.inner = axis == g.dim - 1 ? 1 : g.st[axis];
  a.outer = g.count / (static_cast<int64_t>(a.n) * a.inner);
  int32_t P = std::max<int32_t>(1, kLineMax / a.n);
  P = std::min<int32_t>(P, 16);
  if (a.inner > 1) P = static_cast<int32_t>(std::min<int64_t>(P, a.inner));
  else P = static_cast<int32_t>(std::min<int64_t>(P, a.outer));
  a.P = P;
  a.tiles = a.inner == 1 ? (a.outer + P - 1) / P : a.outer * ((a.inner + P - 1) / P);
  a.rungs = rungs;
  a.F = F;
  a.lds_stride = (a.n + 1) | 1;
  (void)box;
  return a;
}

template <class T>
size_t box_lds(const LineT& a) {
  return static_cast<size_t>(a.P) * (a.lds_stride + 1) * sizeof(double) +
         static_cast<size_t>(a.P) * a.lds_stride * 2 * sizeof(T);
}

// The transposed max / min pass's tile: P lines (the values, the gradient and the sums in slots
// i + i / 32, an odd number of words per line) and per wave the H, G positions and the picks.
constexpr int32_t host_slot32(int32_t i) { return i + (i >> 5); }
template <class T>
size_t ext_lds(const LineT& a, int waves) {
  return static_cast<size_t>(3) * a.P * a.lds_stride * sizeof(T) +
         static_cast<size_t>(waves) * (2 * a.scr_stride + a.pk_stride) * sizeof(uint16_t);
}
template <class T>
int ext_shape(LineT& a) {
  const int32_t span = (a.n + kExtTile - 1) / kExtTile * kExtTile;
  a.lds_stride = host_slot32(span) | 1;
  a.scr_stride = (host_slot32(span) + 1) & ~1;
  a.pk_stride = span;
  // rows: 4 adjacent lines, one per wave; columns: 16 adjacent columns (64-byte runs), 8 waves
  int32_t P = a.inner == 1 ? 4 : 16;
  P = static_cast<int32_t>(std::min<int64_t>(P, a.inner == 1 ? a.outer : a.inner));
  int waves = a.inner == 1 ? 4 : 8;
  for (;;) {
    LineT t = a;
    t.P = P;
    if (ext_lds<T>(t, std::min(waves, P)) <= 160 * 1024 || P == 1) break;
    P /= 2;
  }
  a.P = P;
  a.tiles = a.inner == 1 ? (a.outer + P - 1) / P : a.outer * ((a.inner + P - 1) / P);
  return std::max(1, std::min(waves, P));
}

// One transposed pass over all rungs: vals (max / min: the pass's input), gy (the gradient of its
// output, rung stride F) into gx (sum_rungs: the rungs' sum, added when add; else per rung).
template <class T>
void line_transpose(hipStream_t s, const Geom& g, int axis, int kind /* 0 max, 1 min, 3 box */, const T* vals,
                    int64_t vstride, const T* gy, T* gx, const int32_t* win, int32_t rungs, bool sum_rungs, bool add,
                    bool scale, const int32_t* on) {
  const int64_t F = g.batch * g.count;
  LineT a = line_shape<T>(g, axis, rungs, F, kind == 3);
  const int waves = kind == 3 ? kThreads / 64 : ext_shape<T>(a);
  a.vstride = vstride;
  a.sum_rungs = sum_rungs ? 1 : 0;
  a.add = add ? 1 : 0;
  a.scale = scale ? 1 : 0;
  const int64_t wgs = g.batch * a.tiles;
  if (wgs > 0x7fffffff) throw std::invalid_argument("eda: too many workgroups for a transposed pass");
  a.xcd = kind == 3 || a.inner == 1 ? 0 : xcd_chunk(wgs);
  const dim3 grid(static_cast<unsigned>(wgs), sum_rungs ? 1u : static_cast<unsigned>(rungs));
  const auto go = [&](auto k, size_t lds, auto... args) {
    if (lds > 160 * 1024) throw std::invalid_argument("eda: the line is too long for the backward");
    if (lds > 65536)
      EDA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(lds)));
    EDA_GO(k, grid, dim3(64 * waves), lds, s, args...);
  };
  if (kind == 3) go(k_line_boxT<T>, box_lds<T>(a), a, gy, gx, win, on);
  else if (kind == 0) go(k_line_extT<T, false>, ext_lds<T>(a, waves), a, vals, gy, gx, win, on);
  else go(k_line_extT<T, true>, ext_lds<T>(a, waves), a, vals, gy, gx, win, on);
  EDA_LAUNCH("k_line_T");
}

// ---- the element's weights and its gathers, transposed ----------------------------------------------
template <class T>
struct DEv {
  T d, e, len;
};
template <class T>
__device__ __forceinline__ DEv<T> element_dev(T L, T c, T q, T span_m1) {
  DEv<T> r;
  const T f = (T(1) - c) + c / q;
  r.d = L * f;
  r.len = L * (T(1) + span_m1 * c);
  const T diff = r.len - r.d;
  r.e = T(0.5) * (diff > T(0) ? diff : T(0));
  return r;
}

// t of every voxel from its taped lower rung (the forward's k_weights with idx = lo: the same
// arithmetic, so at the taped residual the same bits).
template <class T>
__global__ void __launch_bounds__(kThreads) k_weights_taped(Geom g, T q, T span_m1, const T* __restrict__ L,
                                                            const T* __restrict__ coh, const ElemDev* __restrict__ el,
                                                            const uint8_t* __restrict__ lo, T* __restrict__ t,
                                                            const int32_t* __restrict__ on) {
  __shared__ double s_lw[kMaxRungs];
  const int64_t b = blockIdx.y;
  if (on != nullptr && on[b] == 0) return;
  const int K = el[b].n_rungs;
  if (threadIdx.x < kMaxRungs)
    s_lw[threadIdx.x] = threadIdx.x < static_cast<unsigned>(K) ? log(static_cast<double>(el[b].rung_w[threadIdx.x])) : 0.0;
  __syncthreads();
  const double w_lo = static_cast<double>(el[b].rung_w[0]);
  const double w_hi = static_cast<double>(el[b].rung_w[K > 0 ? K - 1 : 0]);
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kThreads) + threadIdx.x; i < g.count;
       i += static_cast<int64_t>(gridDim.x) * kThreads) {
    const int64_t at = b * g.count + i;
    T tt = T(0);
    if (K > 1) {
      const T c = coh != nullptr ? coh[at] : T(0);
      const DEv<T> de = element_dev<T>(L[at], c, q, span_m1);
      double dc = static_cast<double>(de.d);
      dc = dc > w_lo ? dc : w_lo;
      dc = dc < w_hi ? dc : w_hi;
      const double lg = log(dc);
      const int idx = lo[at];
      tt = static_cast<T>((lg - s_lw[idx]) / (s_lw[idx + 1] - s_lw[idx]));
    }
    t[at] = tt;
  }
}

// The gradients of t and o (D >= 2) into those of c and n, per voxel.
template <class T>
__global__ void __launch_bounds__(kThreads) k_element_vjp(Geom g, T q, T span_m1, const T* __restrict__ L,
                                                          const T* __restrict__ coh, const T* __restrict__ nvec,
                                                          const ElemDev* __restrict__ el,
                                                          const uint8_t* __restrict__ lo, const T* __restrict__ gt,
                                                          const T* __restrict__ go, T* __restrict__ gc,
                                                          T* __restrict__ gn, const int32_t* __restrict__ on) {
  const int64_t b = blockIdx.y;
  if (on != nullptr && on[b] == 0) return;
  const int64_t F = g.batch * g.count;
  const int K = el[b].n_rungs;
  const double w_lo = static_cast<double>(el[b].rung_w[0]);
  const double w_hi = static_cast<double>(el[b].rung_w[K > 0 ? K - 1 : 0]);
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kThreads) + threadIdx.x; i < g.count;
       i += static_cast<int64_t>(gridDim.x) * kThreads) {
    const int64_t at = b * g.count + i;
    const T c = coh[at];
    const T Lv = L[at];
    const DEv<T> de = element_dev<T>(Lv, c, q, span_m1);
    double ge = 0.0;
    for (int a = 0; a < g.dim; ++a) {
      const double gov = static_cast<double>(go[a * F + at]);
      ge += gov * static_cast<double>(nvec[a * F + at]);
      gn[a * F + at] = static_cast<T>(gov * static_cast<double>(de.e));
    }
    double gd = 0.0, glen = 0.0;
    const double d = static_cast<double>(de.d);
    if (K > 1 && d > w_lo && d < w_hi) {
      const int idx = lo[at];
      const double den = log(static_cast<double>(el[b].rung_w[idx + 1])) - log(static_cast<double>(el[b].rung_w[idx]));
      gd += static_cast<double>(gt[at]) / (d * den);
    }
    if (de.len - de.d > T(0)) {
      gd -= 0.5 * ge;
      glen = 0.5 * ge;
    }
    const double Ld = static_cast<double>(Lv);
    const double qd = static_cast<double>(q);
    gc[at] = static_cast<T>(gd * Ld * (-1.0 + 1.0 / qd) + glen * Ld * static_cast<double>(span_m1));
  }
}

// Half-sample reflected periodic index (as the forward gather).
__device__ __forceinline__ int32_t refl32(int32_t i, int32_t n) {
  if (static_cast<uint32_t>(i) < static_cast<uint32_t>(n)) return i;
  const int32_t p = 2 * n;
  int32_t m = i % p;
  m += m < 0 ? p : 0;
  return m < n ? m : p - 1 - m;
}

// The D-linear interpolation of the forward gather (the same corner offsets and weights), with its
// transpose (scatter) and its gradient in the coordinates.
template <class T, int D>
struct Lerp {
  int32_t off0[D], off1[D];
  T w0[D], w1[D];
  __device__ __forceinline__ void setup(const int32_t* n, const int32_t* st, const double* y) {
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const double fl = floor(y[a]);
      const double tt = y[a] - fl;
      const int32_t i0 = static_cast<int32_t>(fmin(fmax(fl, -1.0e9), 1.0e9));
      const T t1 = static_cast<T>(tt);
      w1[a] = t1;
      w0[a] = T(1) - t1;
      off0[a] = refl32(i0, n[a]) * st[a];
      off1[a] = refl32(i0 + 1, n[a]) * st[a];
    }
  }
  __device__ __forceinline__ int32_t off(int corner) const {
    int32_t o = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) o += ((corner >> (D - 1 - a)) & 1) ? off1[a] : off0[a];
    return o;
  }
  __device__ __forceinline__ T weight(int corner) const {
    T v = T(1);
#pragma unroll
    for (int a = 0; a < D; ++a) v = v * (((corner >> (D - 1 - a)) & 1) ? w1[a] : w0[a]);
    return v;
  }
  __device__ __forceinline__ T at(const T* __restrict__ f) const {
    T acc = T(0);
#pragma unroll
    for (int c = 0; c < (1 << D); ++c) acc = acc + f[off(c)] * weight(c);
    return acc;
  }
  __device__ __forceinline__ void scatter(T* __restrict__ gf, T v) const {
#pragma unroll
    for (int c = 0; c < (1 << D); ++c) unsafeAtomicAdd(gf + off(c), v * weight(c));
  }
  // The scatter with the corners of adjacent lanes merged: where lane l - 1's upper corners along
  // the last axis are lane l's lower ones (the same samples of the same field `tag`: neighbouring
  // voxels whose points moved alike), lane l adds lane l - 1's share to its own lower corners and
  // lane l - 1 leaves its upper ones out, so a run of such lanes makes half the atomics. ok: this
  // lane scatters. Every lane of the wave calls it together (the DPP exchanges read the neighbours'
  // registers); lanes with ok = false neither give nor take.
  __device__ __forceinline__ void scatter_merged(T* __restrict__ gf, T v, bool ok, int32_t tag) const {
    constexpr int DL = D - 1;  // the last axis: corner bit 0
    const int32_t p_tag = dpp_i32<kDppWaveShr1>(-1, ok ? tag : -1);
    const int32_t p_hi = dpp_i32<kDppWaveShr1>(-1, off1[DL]);
    bool same = ok && p_tag == tag && p_hi == off0[DL];
    T pw0[D], pw1[D];
#pragma unroll
    for (int a = 0; a < DL; ++a) {
      const int32_t q0 = dpp_i32<kDppWaveShr1>(-1, off0[a]);
      const int32_t q1 = dpp_i32<kDppWaveShr1>(-1, off1[a]);
      same = same && q0 == off0[a] && q1 == off1[a];
      pw0[a] = dpp_any<kDppWaveShr1>(T(0), w0[a]);
      pw1[a] = dpp_any<kDppWaveShr1>(T(0), w1[a]);
    }
    const T pu = dpp_any<kDppWaveShr1>(T(0), v * w1[DL]);
    const bool gave = dpp_i32<kDppWaveShl1>(0, same ? 1 : 0) != 0;
    if (!ok) return;
#pragma unroll
    for (int c = 0; c < (1 << D); ++c) {
      const bool upper = (c & 1) != 0;
      if (upper && gave) continue;
      T val = v * weight(c);
      if (!upper && same) {
        T wr = pu;
#pragma unroll
        for (int a = 0; a < DL; ++a) wr = wr * (((c >> (D - 1 - a)) & 1) ? pw1[a] : pw0[a]);
        val = val + wr;
      }
      unsafeAtomicAdd(gf + off(c), val);
    }
  }
  // the corner values (one read each), the interpolant and its gradient from them
  __device__ __forceinline__ void corners(const T* __restrict__ f, T* cv) const {
#pragma unroll
    for (int c = 0; c < (1 << D); ++c) cv[c] = f[off(c)];
  }
  __device__ __forceinline__ T at_v(const T* cv) const {
    T acc = T(0);
#pragma unroll
    for (int c = 0; c < (1 << D); ++c) acc = acc + cv[c] * weight(c);
    return acc;
  }
  __device__ __forceinline__ void grad_v(const T* cv, T* dy) const {
#pragma unroll
    for (int a = 0; a < D; ++a) dy[a] = T(0);
#pragma unroll
    for (int c = 0; c < (1 << D); ++c) {
      const T v = cv[c];
#pragma unroll
      for (int a = 0; a < D; ++a) {
        T w = ((c >> (D - 1 - a)) & 1) ? v : -v;
#pragma unroll
        for (int bb = 0; bb < D; ++bb)
          if (bb != a) w = w * (((c >> (D - 1 - bb)) & 1) ? w1[bb] : w0[bb]);
        dy[a] = dy[a] + w;
      }
    }
  }
  // d/dy_a of the interpolant
  __device__ __forceinline__ void grad(const T* __restrict__ f, T* dy) const {
#pragma unroll
    for (int a = 0; a < D; ++a) dy[a] = T(0);
#pragma unroll
    for (int c = 0; c < (1 << D); ++c) {
      const T v = f[off(c)];
#pragma unroll
      for (int a = 0; a < D; ++a) {
        T w = ((c >> (D - 1 - a)) & 1) ? v : -v;
#pragma unroll
        for (int bb = 0; bb < D; ++bb)
          if (bb != a) w = w * (((c >> (D - 1 - bb)) & 1) ? w1[bb] : w0[bb]);
        dy[a] = dy[a] + w;
      }
    }
  }
};

// The gather of one element operator, transposed (OP 0 max, 2 max + min, 3 box). Per voxel and
// each of its two rungs with a positive weight: the rung's value at the voxel and (when the rung
// carries segment points) at the 4 points x + s o, s = 1/2, -1/2, 1, -1; max / min take the first
// point that attains the extreme (the forward's max_c / min_c order) and send the weighted gradient
// there; the box sends a fifth of it to each point. Interpolated points scatter over their corners
// (atomics: points of many voxels land on one sample). gt += g (acc_up - acc_lo) where both rungs
// count; go += the gradient times s times the interpolant's slope at the winning point(s).
template <class T, int D, int OP>
__global__ void __launch_bounds__(kThreads) k_gather_vjp(Geom g, const ElemDev* __restrict__ el,
                                                         const uint8_t* __restrict__ lo, const T* __restrict__ t,
                                                         const T* __restrict__ o, const T* __restrict__ r0,
                                                         const T* __restrict__ r1, const T* __restrict__ g0,
                                                         const T* __restrict__ g1, T* __restrict__ gr0,
                                                         T* __restrict__ gr1, T* __restrict__ gt, T* __restrict__ go,
                                                         const int32_t* __restrict__ on) {
  const int64_t b = blockIdx.y;
  if (on != nullptr && on[b] == 0) return;
  const uint32_t pts = D >= 2 ? el[b].pts : 0u;
  const int64_t base = b * g.count;
  const int64_t F = g.batch * g.count;
  const int32_t count = static_cast<int32_t>(g.count);
  int32_t nn[D], st[D];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    nn[a] = g.n[a];
    st[a] = static_cast<int32_t>(g.st[a]);
  }
  const T sgn[4] = {T(0.5), T(-0.5), T(1), T(-1)};
  const int lane = threadIdx.x & 63;
  // every lane of a wave runs every step of the loop (past the last voxel with valid = false), so
  // that the merged scatters' DPP exchanges read live neighbours
  for (int64_t w0 = blockIdx.x * static_cast<int64_t>(kThreads) + (threadIdx.x & ~63); w0 < count;
       w0 += static_cast<int64_t>(gridDim.x) * kThreads) {
    const bool valid = w0 + lane < count;
    const int32_t i = static_cast<int32_t>(valid ? w0 + lane : w0);
    const int64_t at = base + i;
    const int l = lo[at];
    const T tt = t[at];
    const T wl = T(1) - tt;
    const bool use[2] = {valid && wl > T(0), valid && tt > T(0)};
    const T wgt[2] = {wl, tt};
    const T gv0 = g0[at];
    const T gv1 = OP == 2 ? g1[at] : T(0);
    Lerp<T, D> ip[4] = {};  // defined in every lane (the merged scatters exchange them)
    const bool any_pts = D >= 2 && ((use[0] && ((pts >> l) & 1u)) || (use[1] && ((pts >> (l + 1)) & 1u)));
    if (any_pts) {
      int32_t x[D];
      if constexpr (D == 1) {
        x[0] = i;
      } else if constexpr (D == 2) {
        x[0] = static_cast<int32_t>(static_cast<uint32_t>(i) / static_cast<uint32_t>(nn[1]));
        x[1] = i - x[0] * nn[1];
      } else {
        x[0] = static_cast<int32_t>(static_cast<uint32_t>(i) / static_cast<uint32_t>(st[0]));
        const int32_t rem = i - x[0] * st[0];
        x[1] = static_cast<int32_t>(static_cast<uint32_t>(rem) / static_cast<uint32_t>(nn[2]));
        x[2] = rem - x[1] * nn[2];
      }
      double ov[D];
#pragma unroll
      for (int a = 0; a < D; ++a) ov[a] = static_cast<double>(o[a * F + at]);
#pragma unroll
      for (int k = 1; k <= 2; ++k) {
        const double sc = static_cast<double>(k) / 2.0;
#pragma unroll
        for (int sg = 0; sg < 2; ++sg) {
          double y[D];
#pragma unroll
          for (int a = 0; a < D; ++a) {
            const double off = sc * ov[a];
            const double xa = static_cast<double>(x[a]);
            y[a] = sg == 0 ? xa + off : xa - off;
          }
          ip[2 * (k - 1) + sg].setup(nn, st, y);
        }
      }
    }
    T accv[2][2] = {{T(0), T(0)}, {T(0), T(0)}};  // [side][max, min]
    T god[D];
#pragma unroll
    for (int a = 0; a < D; ++a) god[a] = T(0);
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const bool us = use[side];
      const int j = l + side;
      const bool p = us && D >= 2 && ((pts >> j) & 1u);
      const T* C0 = r0 + j * F + base;
      T* G0 = gr0 + j * F + base;
      const T wj = wgt[side];
      if constexpr (OP == 3) {
        T acc = T(0);
        T cv[4][1 << D];  // the points' corner values, read once for the value and the slope
        if (us) {
          acc = C0[i];
          if (p) {
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) {
              ip[qq].corners(C0, cv[qq]);
              acc = acc + ip[qq].at_v(cv[qq]);
            }
            acc = acc / T(5);
          }
        }
        accv[side][0] = acc;
        const T gs = wj * gv0;
        const T g5 = gs / T(5);
        if (us) unsafeAtomicAdd(G0 + i, p ? g5 : gs);
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) {
          ip[qq].scatter_merged(G0, g5, p, j);
          if (p && go != nullptr) {
            T dy[D];
            ip[qq].grad_v(cv[qq], dy);
#pragma unroll
            for (int a = 0; a < D; ++a) god[a] = god[a] + g5 * sgn[qq] * dy[a];
          }
        }
      } else {
#pragma unroll
        for (int part = 0; part < (OP == 2 ? 2 : 1); ++part) {
          const bool mn = part == 1;
          const T* C = mn ? r1 + j * F + base : C0;
          T* G = mn ? gr1 + j * F + base : G0;
          T best = T(0);
          int win = -1;
          T wcv[1 << D];  // the winning point's corner values (its slope)
          if (us) {
            best = C[i];
            if (p) {
#pragma unroll
              for (int qq = 0; qq < 4; ++qq) {
                T cv[1 << D];
                ip[qq].corners(C, cv);
                const T v = ip[qq].at_v(cv);
                const bool keep = mn ? (best < v || (best == v && signbit(best))) : (best > v || (best == v && !signbit(best)));
                if (!keep) {
                  best = v;
                  win = qq;
#pragma unroll
                  for (int c = 0; c < (1 << D); ++c) wcv[c] = cv[c];
                }
              }
            }
          }
          accv[side][part] = best;
          const T gs = wj * (mn ? gv1 : gv0);
          if (us && win < 0) unsafeAtomicAdd(G + i, gs);
          // the winning point (selected without indexing the register array at run time)
          Lerp<T, D> wp = ip[0];
          if (win == 1) wp = ip[1];
          if (win == 2) wp = ip[2];
          if (win == 3) wp = ip[3];
          wp.scatter_merged(G, gs, us && win >= 0, 2 * j + part);
          if (us && win >= 0 && go != nullptr) {
            T dy[D];
            wp.grad_v(wcv, dy);
#pragma unroll
            for (int a = 0; a < D; ++a) god[a] = god[a] + gs * sgn[win] * dy[a];
          }
        }
      }
    }
    if (use[0] && use[1] && gt != nullptr) {
      T dgt = gv0 * (accv[1][0] - accv[0][0]);
      if constexpr (OP == 2) dgt = dgt + gv1 * (accv[1][1] - accv[0][1]);
      gt[at] = gt[at] + dgt;
    }
    if (go != nullptr && any_pts) {
#pragma unroll
      for (int a = 0; a < D; ++a) go[a * F + at] = go[a * F + at] + god[a];
    }
  }
}

dim3 elem_grid_t(const Geom& g) {
  int64_t bx = (g.count + kThreads - 1) / kThreads;
  const int64_t want = std::max<int64_t>(16, (int64_t{8192} + g.batch - 1) / g.batch);
  bx = std::min(bx, std::min<int64_t>(1024, want));
  bx = bx < 1 ? 1 : bx;
  return dim3(static_cast<unsigned>(bx), static_cast<unsigned>(g.batch));
}

template <class T, int OP>
void launch_gather_vjp(hipStream_t s, const Geom& g, const ElemDev* el, const uint8_t* lo, const T* t, const T* o,
                       const T* r0, const T* r1, const T* g0, const T* g1, T* gr0, T* gr1, T* gt, T* go,
                       const int32_t* on) {
  const dim3 grid = elem_grid_t(g);
  if (g.dim == 1)
    EDA_GO((k_gather_vjp<T, 1, OP>), grid, dim3(kThreads), 0, s, g, el, lo, t, o, r0, r1, g0, g1, gr0, gr1, gt,
           static_cast<T*>(nullptr), on);
  else if (g.dim == 2)
    EDA_GO((k_gather_vjp<T, 2, OP>), grid, dim3(kThreads), 0, s, g, el, lo, t, o, r0, r1, g0, g1, gr0, gr1, gt, go,
           on);
  else
    EDA_GO((k_gather_vjp<T, 3, OP>), grid, dim3(kThreads), 0, s, g, el, lo, t, o, r0, r1, g0, g1, gr0, gr1, gt, go,
           on);
  EDA_LAUNCH("k_gather_vjp");
}


// ---- the backward driver -------------------------------------------------------------------------
// Device copies of a mode's per-element parameters (uploaded once per call, indexed by mode).
struct ModeDev {
  int32_t* on = nullptr;       // [B] active
  int32_t* w = nullptr;        // [B]
  int32_t* W = nullptr;        // [B]
  double* tol = nullptr;       // [B]
  ElemDev* el = nullptr;       // [B]
  int32_t* rw = nullptr;       // [K_max B]
  std::vector<int32_t*> live;  // per sift j >= 1: [B] ran and accepted
  std::vector<int32_t*> rwj;   // per sift: [K_max B], the rungs of the live elements
  std::vector<double*> thr;    // per sift: [B]
  std::vector<int32_t> rw_max; // per sift: the largest rung side of each rung over its live elements
  std::vector<int32_t> any_live;
  int32_t mW = 0;
};

template <class T>
class Adjoint {
 public:
  Adjoint(hipStream_t s, int device, const MethodCtl& c, const ChunkTape& ct, bool orient_grad)
      : s_(s), dev_(device), c_(c), ct_(ct), g_(ct.g), orient_grad_(orient_grad && ct.g.dim >= 2) {
    if (c_.anorm_smooth != 1 || c_.u != 2 || c_.s != 1)
      throw std::logic_error("eda: the backward supports the default smoothing counts only");
    for (int a = 0; a < g_.dim; ++a)
      if (g_.n[a] > kLineMax)
        throw std::invalid_argument("eda: the backward takes axes of at most " + std::to_string(kLineMax) +
                                    " samples");
    B_ = g_.batch;
    F_ = B_ * g_.count;
    const int D = g_.dim;
    const int K = D * (D + 1) / 2;
    sifts_ = 0;
    K_max_ = 1;
    for (const TapeMode& m : ct.modes) {
      sifts_ = std::max<int>(sifts_, static_cast<int>(m.ran.size()));
      K_max_ = std::max(K_max_, m.K_max);
    }
    // fields: orientation (n, c, o, t, work), sift 1 and the continuation weight, per sift, gradients
    nsift_f_ = 18;
    const int fixed = D + 1 + D + 1 + (D + K + 4) + 11 + sifts_ * nsift_f_ + 1 + 5 + kScratch + 2 * D + 1;
    const int rung = (6 + std::max(1, D - 1)) * K_max_;
    const size_t fb = static_cast<size_t>(F_) * sizeof(T);
    const size_t small = static_cast<size_t>(orient_scratch_doubles(B_)) * 8 + static_cast<size_t>(B_) * 64 + 4096;
    mem_ = TapeBuf(static_cast<size_t>(fixed + rung) * (fb + 256) + small);
    char* base = static_cast<char*>(mem_.p);
    size_t off = 0;
    auto take = [&](size_t bytes) {
      off = (off + 255) / 256 * 256;
      char* q = base + off;
      off += bytes;
      return q;
    };
    auto field = [&](int k = 1) { return reinterpret_cast<T*>(take(fb * k)); };
    nvec_ = field(D);
    coh_ = field();
    o_ = field(D);
    t_ = field();
    owork_ = field(D + K + 4);
    u1_ = field(); l1_ = field(); mid1_ = field(); R_ = field(); h1_ = field(); V1_ = field(); V2_ = field();
    gw_ = field(); A0_ = field(); a_ = field(); m1_ = field();
    sf_.resize(sifts_);
    for (int j = 0; j < sifts_; ++j) {
      sf_[j].resize(nsift_f_);
      for (int k = 0; k < nsift_f_; ++k) sf_[j][k] = field();
    }
    hlast_ = field();
    gh_ = field(); ghj_ = field(); gr_ = field(); ga_ = field(); gt_ = field();
    for (int k = 0; k < kScratch; ++k) gs_[k] = field();
    go_ = field(D);
    gn_ = field(D);
    gc_ = field();
    r0_ = field(K_max_); r1_ = field(K_max_); gC0_ = field(K_max_); gC1_ = field(K_max_);
    Z_ = field(std::max(1, D - 1) * K_max_);
    G_ = field(2 * K_max_);
    dscr_ = reinterpret_cast<double*>(take(static_cast<size_t>(orient_scratch_doubles(B_)) * 8));
    dmm_ = reinterpret_cast<uint64_t*>(take(static_cast<size_t>(B_) * 16));
    nonfin_ = reinterpret_cast<int32_t*>(take(static_cast<size_t>(B_) * 4));
    if (off > static_cast<size_t>(fixed + rung) * (fb + 256) + small) throw std::logic_error("eda: adjoint layout");
    upload_params();
  }

  // grad_x (B, count) = the gradient of <grad_rows, rows> for grad_rows (B, P, count) of this chunk.
  void backward(const T* grad_rows, int32_t P, T* grad_x) {
    const Geom g = g_;
    const int64_t cnt = g.count;
    T* gr = gr_;
    const int32_t* kx = kexp_;
    const int32_t* nim = nimfs_dev_;
    // the residue row's gradient, scaled to the prescaled residual
    each(s_, g, nullptr, [=] __device__(int64_t b, int64_t i) {
      const int32_t k = nim[b];
      gr[b * cnt + i] = k < 0 ? T(0) : static_cast<T>(ldexp(static_cast<double>(grad_rows[(b * P + k) * cnt + i]), kx[b]));
    });
    for (int m = static_cast<int>(ct_.modes.size()) - 1; m >= 0; --m) {
      const ModeDev& md = md_[m];
      T* gh = gh_;
      const int32_t row = m;
      // gh = grad_rows[m] 2^k - gr (the mode's own gradient minus what flows on through r' = r - h)
      each(s_, g, md.on, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * cnt + i;
        const T gv = static_cast<T>(ldexp(static_cast<double>(grad_rows[(b * P + row) * cnt + i]), kx[b]));
        gh[at] = gv - gr[at];
      });
      mode_vjp(m, static_cast<const T*>(ct_.modes[m].r.p));
    }
    each(s_, g, nullptr, [=] __device__(int64_t b, int64_t i) {
      const int64_t at = b * cnt + i;
      grad_x[at] = nim[b] < 0 ? T(0) : static_cast<T>(ldexp(static_cast<double>(gr[at]), -kx[b]));
    });
  }

  // rows (B, P, count): the frozen map's rows at x (B, count), zero above an element's n_imfs.
  void replay(const T* x, T* rows, int32_t P) {
    const Geom g = g_;
    const int64_t cnt = g.count;
    const int32_t* kx = kexp_;
    const int32_t* nim = nimfs_dev_;
    T* r = gr_;  // the running residual (the gradient field is free in a replay)
    EDA_HIP(hipMemsetAsync(rows, 0, sizeof(T) * static_cast<size_t>(B_) * P * cnt, s_));
    each(s_, g, nullptr, [=] __device__(int64_t b, int64_t i) {
      const int64_t at = b * cnt + i;
      if constexpr (std::is_same_v<T, float>) r[at] = static_cast<float>(ldexp(static_cast<double>(x[at]), -kx[b]));
      else r[at] = ldexp(x[at], -kx[b]);
    });
    for (int m = 0; m < static_cast<int>(ct_.modes.size()); ++m) {
      const T* h = mode_forward(m, r);
      const int32_t row = m;
      each(s_, g, md_[m].on, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * cnt + i;
        const T hv = h[at];
        T v;
        if constexpr (std::is_same_v<T, float>) v = static_cast<float>(ldexp(static_cast<double>(hv), kx[b]));
        else v = ldexp(hv, kx[b]);
        rows[(b * P + row) * cnt + i] = v;
        r[at] = r[at] - hv;
      });
    }
    each(s_, g, nullptr, [=] __device__(int64_t b, int64_t i) {
      const int32_t k = nim[b];
      if (k < 0) return;
      T v;
      if constexpr (std::is_same_v<T, float>) v = static_cast<float>(ldexp(static_cast<double>(r[b * cnt + i]), kx[b]));
      else v = ldexp(r[b * cnt + i], kx[b]);
      rows[(b * P + k) * cnt + i] = v;
    });
  }

 private:
  enum SiftField { kH = 0, kUp, kLo, kMid, kPlain, kHd, kE0, kEmax, kEmin, kKr, kGk0, kE1, kQ, kUq, kLq, kMidq, kMn0, kHn };

  void upload_params() {
    // every mode's per-element arrays in one host block, copied once
    const int D = g_.dim;
    (void)D;
    std::vector<char> host;
    std::vector<std::pair<size_t, void**>> fix;  // offsets to patch into device pointers
    auto put = [&](const void* data, size_t bytes, void** dst) {
      const size_t at = (host.size() + 255) / 256 * 256;
      host.resize(at + bytes);
      std::memcpy(host.data() + at, data, bytes);
      fix.emplace_back(at, dst);
    };
    md_.resize(ct_.modes.size());
    std::vector<int32_t> kexp = ct_.kexp, nim(B_);
    for (int64_t b = 0; b < B_; ++b) nim[b] = ct_.error[b] ? -1 : ct_.n_imfs[b];
    for (size_t m = 0; m < ct_.modes.size(); ++m) {
      const TapeMode& tm = ct_.modes[m];
      ModeDev& md = md_[m];
      put(tm.active.data(), B_ * 4, reinterpret_cast<void**>(&md.on));
      put(tm.w.data(), B_ * 4, reinterpret_cast<void**>(&md.w));
      put(tm.W.data(), B_ * 4, reinterpret_cast<void**>(&md.W));
      put(tm.tol.data(), B_ * 8, reinterpret_cast<void**>(&md.tol));
      put(tm.el.data(), B_ * sizeof(ElemDev), reinterpret_cast<void**>(&md.el));
      std::vector<int32_t> rw(static_cast<size_t>(K_max_) * B_, 0);
      for (int j = 0; j < tm.K_max; ++j)
        for (int64_t b = 0; b < B_; ++b) rw[j * B_ + b] = tm.rung_win[j * B_ + b];
      put(rw.data(), rw.size() * 4, reinterpret_cast<void**>(&md.rw));
      md.mW = 0;
      for (int64_t b = 0; b < B_; ++b)
        if (tm.active[b]) md.mW = std::max(md.mW, tm.W[b]);
      const int ns = static_cast<int>(tm.ran.size());
      md.live.assign(ns, nullptr);
      md.rwj.assign(ns, nullptr);
      md.thr.assign(ns, nullptr);
      md.any_live.assign(ns, 0);
      for (int j = 0; j < ns; ++j) {
        std::vector<int32_t> lv(B_, 0), rj(static_cast<size_t>(K_max_) * B_, 0);
        for (int64_t b = 0; b < B_; ++b) {
          lv[b] = tm.active[b] && tm.ran[j][b] && tm.acc[j][b] ? 1 : 0;
          md.any_live[j] |= lv[b];
          for (int k = 0; k < tm.K_max; ++k) rj[k * B_ + b] = lv[b] ? tm.rung_win[k * B_ + b] : 0;
        }
        put(lv.data(), B_ * 4, reinterpret_cast<void**>(&md.live[j]));
        put(rj.data(), rj.size() * 4, reinterpret_cast<void**>(&md.rwj[j]));
        put(tm.thr[j].data(), B_ * 8, reinterpret_cast<void**>(&md.thr[j]));
      }
    }
    put(kexp.data(), B_ * 4, reinterpret_cast<void**>(&kexp_));
    put(nim.data(), B_ * 4, reinterpret_cast<void**>(&nimfs_dev_));
    params_ = TapeBuf(host.size() + 256);
    EDA_HIP(hipMemcpyAsync(params_.p, host.data(), host.size(), hipMemcpyHostToDevice, s_));
    for (auto& f : fix) *f.second = static_cast<char*>(params_.p) + f.first;
    EDA_HIP(hipStreamSynchronize(s_));  // the host block is freed on return
  }

  // The forward operator `op` (0 Dil, 2 Dil + Ero, 3 Avg) of mode m with the rung windows rw.
  void op_fwd(const ModeDev& md, int m, int op, const T* in, T* out0, T* out1, const int32_t* on, const int32_t* rw,
              const GatherExtra<T>& extra = {}) {
    const TapeMode& tm = ct_.modes[m];
    if (tm.K_max < 1) return;
    enqueue_element_op<T>(s_, g_, op, md.el, lo_, t_, o_, rw, tm.rung_max.data(), tm.K_max, in, out0, out1, r0_, r1_,
                          nullptr, on, extra);
  }

  // gin += the operator's transpose applied to (g0, g1) at the input `in`; gt, go collect the
  // gradients of the hat weight and the segment vector.
  void op_vjp(const ModeDev& md, int m, int op, const T* in, const T* g0, const T* g1, T* gin, const int32_t* on,
              const int32_t* rw) {
    const TapeMode& tm = ct_.modes[m];
    const int Km = tm.K_max;
    if (Km < 1) return;
    int32_t mw = 0;
    for (int j = 0; j < Km; ++j) mw = std::max(mw, tm.rung_max[j]);
    const CubeOp cop = op == 3 ? CubeOp::kBox : (op == 2 ? CubeOp::kMaxMin : CubeOp::kMax);
    enqueue_cube_rungs<T>(s_, g_, cop, in, r0_, r1_, rw, Km, F_, mw, nullptr, op == 2 ? nullptr : r1_);
    EDA_HIP(hipMemsetAsync(gC0_, 0, sizeof(T) * static_cast<size_t>(Km) * F_, s_));
    if (op == 2) EDA_HIP(hipMemsetAsync(gC1_, 0, sizeof(T) * static_cast<size_t>(Km) * F_, s_));
    T* go = orient_grad_ ? go_ : nullptr;
    if (op == 3)
      launch_gather_vjp<T, 3>(s_, g_, md.el, lo_, t_, o_, r0_, r1_, g0, g1, gC0_, gC1_, gt_, go, on);
    else if (op == 2)
      launch_gather_vjp<T, 2>(s_, g_, md.el, lo_, t_, o_, r0_, r1_, g0, g1, gC0_, gC1_, gt_, go, on);
    else
      launch_gather_vjp<T, 0>(s_, g_, md.el, lo_, t_, o_, r0_, r1_, g0, g1, gC0_, gC1_, gt_, go, on);
    cube_vjp(op == 3 ? 3 : 0, in, gC0_, gin, rw, Km, mw, on);
    if (op == 2) cube_vjp(1, in, gC1_, gin, rw, Km, mw, on);
  }

  // gf += the transpose of the separable cube (kind 0 max, 1 min, 3 box) of `rungs` windows at f.
  void cube_vjp(int kind, const T* f, const T* gC, T* gf, const int32_t* win, int32_t rungs, int32_t mw,
                const int32_t* on) {
    const int D = g_.dim;
    if (kind == 3) {
      // forward: axes 0 .. D - 1, the last scaled; transposed: D - 1 .. 0
      const T* src = gC;
      for (int a = D - 1; a >= 0; --a) {
        const bool last = a == 0;
        T* dst = last ? gf : G_ + static_cast<int64_t>((D - 1 - a) & 1) * K_max_ * F_;
        line_transpose<T>(s_, g_, a, 3, nullptr, 0, src, dst, win, rungs, last, last, a == D - 1, on);
        src = dst;
      }
      return;
    }
    // forward order chosen here: axes 0 .. D - 1 (max and min are exact in any order); the
    // intermediate inputs z_1 .. z_{D-1} recomputed
    const T* zin[3] = {f, nullptr, nullptr};
    for (int a = 1; a < D; ++a) {
      T* z = Z_ + static_cast<int64_t>(a - 1) * K_max_ * F_;
      enqueue_axis_pass<T>(s_, g_, a - 1, kind == 1, a == 1 ? f : zin[a - 1], z, win, mw, rungs, a == 1 ? 0 : F_, F_);
      zin[a] = z;
    }
    const T* src = gC;
    for (int a = D - 1; a >= 0; --a) {
      const bool last = a == 0;
      T* dst = last ? gf : G_ + static_cast<int64_t>((D - 1 - a) & 1) * K_max_ * F_;
      line_transpose<T>(s_, g_, a, kind, zin[a], a == 0 ? 0 : F_, src, dst, win, rungs, last, last, false, on);
      src = dst;
    }
  }

  // The mode's forward from r with its taped decisions; returns h (the mode). Fills the fields the
  // backward reads.
  const T* mode_forward(int m, const T* r) {
    const TapeMode& tm = ct_.modes[m];
    const ModeDev& md = md_[m];
    const Geom g = g_;
    const MethodCtl c = c_;
    const int32_t* on = md.on;
    if (g.dim >= 2)
      enqueue_orientation<T>(s_, dev_, g, r, md.w, on, nvec_, coh_, owork_, nonfin_, dscr_);
    const T* L = static_cast<const T*>(tm.L.p);
    lo_ = static_cast<const uint8_t*>(tm.lo.p);
    enqueue_element_build<T>(s_, g, c, L, nvec_, coh_, o_, on, dmm_);
    EDA_GO(k_weights_taped<T>, elem_grid_t(g), dim3(kThreads), 0, s_, g, static_cast<T>(c.q),
           static_cast<T>(c.span - 1.0), L, g.dim >= 2 ? coh_ : nullptr, md.el, lo_, t_, on);
    EDA_LAUNCH("k_weights_taped");
    // sift 1: u, l, mid, R = max(u - l, 0) from the max / min gather; h = r - Avg(mid)
    {
      GatherExtra<T> mm;
      mm.mid = mid1_;
      mm.pdiff = c.sifts > 1 ? R_ : nullptr;
      op_fwd(md, m, 2, r, u1_, l1_, on, md.rw, mm);
      GatherExtra<T> bx;
      bx.minuend = r;
      bx.diff = h1_;
      op_fwd(md, m, 3, mid1_, m1_, nullptr, on, md.rw, bx);
    }
    const double* dtol = md.tol;
    if (c.sifts > 1) {
      op_fwd(md, m, 0, R_, V1_, nullptr, on, md.rw);
      op_fwd(md, m, 0, V1_, V2_, nullptr, on, md.rw);
      const T t0 = static_cast<T>(c.t0);
      const double dd = c.t1 - c.t0;
      const T den = static_cast<T>(dd >= 1e-12 ? dd : 1e-12);
      const T* Uc = V2_;
      const T* R = R_;
      T* gw = gw_;
      each(s_, g, on, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        const T Uv = Uc[at];
        T rho = T(0);
        if (static_cast<double>(Uv) > dtol[b]) rho = clip_np<T>(R[at] / Uv, T(0), T(1));
        gw[at] = clip_np<T>((rho - t0) / den, T(0), T(1));
      });
      op_fwd(md, m, 3, gw_, A0_, nullptr, on, md.rw);
      const T* A0 = A0_;
      T* a = a_;
      each(s_, g, on, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        a[at] = clip_np<T>(A0[at], T(0), T(1));
      });
    }
    const T* h = h1_;
    for (int j = 0; j < static_cast<int>(tm.ran.size()); ++j) {
      if (!md.any_live[j]) break;
      std::vector<T*>& f = sf_[j];
      const int32_t* son = md.live[j];
      const int32_t* rwj = md.rwj[j];
      // the sift's input, kept for the backward
      if (h != f[kH]) EDA_HIP(hipMemcpyAsync(f[kH], h, sizeof(T) * F_, hipMemcpyDeviceToDevice, s_));
      const T* hj = f[kH];
      {
        GatherExtra<T> mm;
        mm.mid = f[kMid];
        mm.hdiff = f[kHd];
        op_fwd(md, m, 2, hj, f[kUp], f[kLo], son, rwj, mm);
        GatherExtra<T> bx;
        bx.minuend = hj;
        op_fwd(md, m, 3, f[kMid], f[kPlain], nullptr, son, rwj, bx);
      }
      op_fwd(md, m, 3, f[kHd], f[kE0], nullptr, son, rwj);
      const int32_t* dW = md.W;
      enqueue_cube<T>(s_, g, CubeOp::kMaxMin, f[kE0], f[kEmax], f[kEmin], masked_W(m, j), md.mW);
      {
        const T k0 = static_cast<T>(c.ramp_k0);
        const double dd = c.ramp_k1 - c.ramp_k0;
        const T kden = static_cast<T>(dd >= 1e-12 ? dd : 1e-12);
        const double tiny = c.tiny;
        const T* emax = f[kEmax];
        const T* emin = f[kEmin];
        T* kr = f[kKr];
        each(s_, g, son, [=] __device__(int64_t b, int64_t i) {
          const int64_t at = b * g.count + i;
          const double tl = dtol[b];
          const T fl = static_cast<T>(tl >= tiny ? tl : tiny);
          const T hi = emax[at], lw = emin[at];
          const T sum = hi + lw;
          const T den = sum >= fl ? sum : fl;
          const T kap = (hi - lw) / den;
          kr[at] = clip_np<T>((kap - k0) / kden, T(0), T(1));
        });
      }
      (void)dW;
      enqueue_cube<T>(s_, g, CubeOp::kBox, f[kKr], f[kGk0], nullptr, masked_W(m, j), md.mW);
      enqueue_cube<T>(s_, g, CubeOp::kBox, f[kE0], f[kE1], nullptr, masked_W(m, j), md.mW);
      {
        const double* thr = md.thr[j];
        const T* e1 = f[kE1];
        T* q = f[kQ];
        each(s_, g, son, [=] __device__(int64_t b, int64_t i) {
          const int64_t at = b * g.count + i;
          const double th = thr[b];
          const T ev = e1[at];
          const T ef = static_cast<double>(ev) >= th ? ev : static_cast<T>(th);
          q[at] = hj[at] / ef;
        });
      }
      {
        GatherExtra<T> mm;
        mm.mid = f[kMidq];
        op_fwd(md, m, 2, f[kQ], f[kUq], f[kLq], son, rwj, mm);
        op_fwd(md, m, 3, f[kMidq], f[kMn0], nullptr, son, rwj);
      }
      {
        const double* thr = md.thr[j];
        const T* gk0 = f[kGk0];
        const T* e1 = f[kE1];
        const T* Mn0 = f[kMn0];
        const T* up = f[kUp];
        const T* lo = f[kLo];
        const T* plain = f[kPlain];
        const T* a = a_;
        T* hn = f[kHn];
        each(s_, g, son, [=] __device__(int64_t b, int64_t i) {
          const int64_t at = b * g.count + i;
          const T wv = clip_np<T>(gk0[at], T(0), T(1));
          const double th = thr[b];
          const T ev = e1[at];
          const T ef = static_cast<double>(ev) >= th ? ev : static_cast<T>(th);
          const T mA = clip_np<T>(ef * Mn0[at], lo[at], up[at]);
          const T bl = wv * mA + (T(1) - wv) * plain[at];
          const T mm = a[at] * bl;
          hn[at] = hj[at] - mm;
        });
      }
      // the next sift's input: hn for the live elements, the same h for the others
      T* hnext = j + 1 < sifts_ ? sf_[j + 1][kH] : hlast_;
      {
        const T* hn = f[kHn];
        const int32_t* lv = son;
        each(s_, g, on, [=] __device__(int64_t b, int64_t i) {
          const int64_t at = b * g.count + i;
          hnext[at] = lv[b] ? hn[at] : hj[at];
        });
      }
      h = hnext;
    }
    return h;
  }

  // W of the elements live in sift j (0 elsewhere): G1's cubes run on them only.
  const int32_t* masked_W(int m, int j) {
    const ModeDev& md = md_[m];
    const int32_t* W = md.W;
    const int32_t* lv = md.live[j];
    int32_t* dst = nonfin_;  // a [B] int32 scratch
    const int64_t B = B_;
    flat(s_, B, [=] __device__(int64_t b) { dst[b] = lv[b] ? W[b] : 0; });
    return dst;
  }

  // gr += the mode's transpose applied to gh_ (the gradient of its h) at the taped residual r.
  void mode_vjp(int m, const T* r) {
    const TapeMode& tm = ct_.modes[m];
    const ModeDev& md = md_[m];
    const Geom g = g_;
    const MethodCtl c = c_;
    const int32_t* on = md.on;
    const int D = g.dim;
    const T* hout = mode_forward(m, r);
    (void)hout;
    EDA_HIP(hipMemsetAsync(gt_, 0, sizeof(T) * F_, s_));
    if (orient_grad_) EDA_HIP(hipMemsetAsync(go_, 0, sizeof(T) * F_ * D, s_));
    EDA_HIP(hipMemsetAsync(ga_, 0, sizeof(T) * F_, s_));
    T* gh = gh_;
    T* ghj = ghj_;
    T* ga = ga_;
    const double* dtol = md.tol;
    const T* a = a_;
    // sifts, last to first (only the elements whose sift j ran and was accepted: for the others h
    // passed through unchanged)
    for (int j = static_cast<int>(tm.ran.size()) - 1; j >= 0; --j) {
      if (!md.any_live[j]) continue;
      std::vector<T*>& f = sf_[j];
      const int32_t* son = md.live[j];
      const int32_t* rwj = md.rwj[j];
      const double* thr = md.thr[j];
      const T* hj = f[kH];
      const T* gk0 = f[kGk0];
      const T* e1 = f[kE1];
      const T* Mn0 = f[kMn0];
      const T* up = f[kUp];
      const T* lo = f[kLo];
      const T* plain = f[kPlain];
      const T* q = f[kQ];
      T* g_gk0 = gs_[0];
      T* g_plain = gs_[1];
      T* g_up = gs_[2];
      T* g_lo = gs_[3];
      T* g_Mn0 = gs_[4];
      T* g_e = gs_[5];
      T* g_midq = gs_[6];
      T* g_q = gs_[7];
      T* g_e0 = gs_[8];
      T* g_kr = gs_[9];
      T* g_emax = gs_[10];
      T* g_emin = gs_[11];
      T* g_hd = gs_[12];
      T* g_mid = gs_[13];
      T* g_half = gs_[14];
      // hn = h - a blend, blend = wv clip(e Mn0, lo, up) + (1 - wv) plain, wv = clip(gk0, 0, 1)
      each(s_, g, son, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        const T ghn = gh[at];
        const T gk0v = gk0[at];
        const T wv = clip_np<T>(gk0v, T(0), T(1));
        const double th = thr[b];
        const T ev = e1[at];
        const T ef = static_cast<double>(ev) >= th ? ev : static_cast<T>(th);
        const T v = ef * Mn0[at];
        const T lov = lo[at], upv = up[at];
        const T mA = clip_np<T>(v, lov, upv);
        const T pl = plain[at];
        const T bl = wv * mA + (T(1) - wv) * pl;
        ga[at] = ga[at] - ghn * bl;
        const T gbl = -ghn * a[at];
        g_gk0[at] = (gk0v >= T(0) && gk0v <= T(1)) ? gbl * (mA - pl) : T(0);
        g_plain[at] = gbl * (T(1) - wv);
        const T gm = gbl * wv;
        T gv = T(0), gl = T(0), gu = T(0);
        if (v < lov) gl = gm;
        else if (v > upv) gu = gm;
        else gv = gm;
        g_up[at] = gu;
        g_lo[at] = gl;
        g_Mn0[at] = gv * ef;
        g_e[at] = gv * Mn0[at];
        ghj[at] = ghn;
      });
      // Mn0 = Avg(midq), midq = 0.5 (uq + lq), (uq, lq) = DilEro(q)
      EDA_HIP(hipMemsetAsync(g_midq, 0, sizeof(T) * F_, s_));
      op_vjp(md, m, 3, f[kMidq], g_Mn0, nullptr, g_midq, son, rwj);
      each(s_, g, son, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        g_half[at] = T(0.5) * g_midq[at];
      });
      EDA_HIP(hipMemsetAsync(g_q, 0, sizeof(T) * F_, s_));
      op_vjp(md, m, 2, q, g_half, g_half, g_q, son, rwj);
      // q = h / e, e = max(e1, thr): h gets gq / e; e1 gets [e1 >= thr] (ge - gq q / e)
      each(s_, g, son, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        const double th = thr[b];
        const T ev = e1[at];
        const bool pass = static_cast<double>(ev) >= th;
        const T ef = pass ? ev : static_cast<T>(th);
        const T gqq = g_q[at];
        ghj[at] = ghj[at] + gqq / ef;
        const T ge = g_e[at] - gqq * q[at] / ef;
        g_e[at] = pass ? ge : T(0);
      });
      // e1 = Box_W(e0), gk0 = Box_W(kr)
      const int32_t* mW = masked_W(m, j);
      EDA_HIP(hipMemsetAsync(g_e0, 0, sizeof(T) * F_, s_));
      cube_vjp(3, nullptr, g_e, g_e0, mW, 1, md.mW, son);
      EDA_HIP(hipMemsetAsync(g_kr, 0, sizeof(T) * F_, s_));
      cube_vjp(3, nullptr, g_gk0, g_kr, mW, 1, md.mW, son);
      // kr = clip((kap - k0) / kden, 0, 1), kap = (emax - emin) / max(emax + emin, fl)
      {
        const T k0 = static_cast<T>(c.ramp_k0);
        const double dd = c.ramp_k1 - c.ramp_k0;
        const T kden = static_cast<T>(dd >= 1e-12 ? dd : 1e-12);
        const double tiny = c.tiny;
        const T* emax = f[kEmax];
        const T* emin = f[kEmin];
        each(s_, g, son, [=] __device__(int64_t b, int64_t i) {
          const int64_t at = b * g.count + i;
          const double tl = dtol[b];
          const T fl = static_cast<T>(tl >= tiny ? tl : tiny);
          const T hi = emax[at], lw = emin[at];
          const T sum = hi + lw;
          const bool sum_in = sum >= fl;
          const T den = sum_in ? sum : fl;
          const T kap = (hi - lw) / den;
          const T u = (kap - k0) / kden;
          const T gk = (u >= T(0) && u <= T(1)) ? g_kr[at] / kden : T(0);
          const T dsum = sum_in ? kap / den : T(0);
          g_emax[at] = gk * (T(1) / den - dsum);
          g_emin[at] = gk * (-(T(1) / den) - dsum);
        });
      }
      cube_vjp(0, f[kE0], g_emax, g_e0, mW, 1, md.mW, son);
      cube_vjp(1, f[kE0], g_emin, g_e0, mW, 1, md.mW, son);
      // e0 = Avg(hd), hd = 0.5 (up - lo); plain = Avg(mid), mid = 0.5 (up + lo)
      EDA_HIP(hipMemsetAsync(g_hd, 0, sizeof(T) * F_, s_));
      op_vjp(md, m, 3, f[kHd], g_e0, nullptr, g_hd, son, rwj);
      EDA_HIP(hipMemsetAsync(g_mid, 0, sizeof(T) * F_, s_));
      op_vjp(md, m, 3, f[kMid], g_plain, nullptr, g_mid, son, rwj);
      each(s_, g, son, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        const T hd = T(0.5) * g_hd[at], md2 = T(0.5) * g_mid[at];
        g_up[at] = g_up[at] + hd + md2;
        g_lo[at] = g_lo[at] - hd + md2;
      });
      op_vjp(md, m, 2, hj, g_up, g_lo, ghj, son, rwj);
      // the live elements' gradient moves to the sift's input
      each(s_, g, son, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        gh[at] = ghj[at];
      });
    }
    T* gA = gs_[0];
    T* gB = gs_[1];
    T* gC = gs_[2];
    T* gD = gs_[3];
    T* gE = gs_[4];
    // the continuation weight: a = clip(Avg(gw)), gw = ramp(R / Dil(Dil(R))), R = max(u1 - l1, 0)
    T* gR = gA;
    EDA_HIP(hipMemsetAsync(gR, 0, sizeof(T) * F_, s_));
    if (c.sifts > 1) {
      const T* A0 = A0_;
      T* gA0 = gB;
      each(s_, g, on, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        const T v = A0[at];
        gA0[at] = (v >= T(0) && v <= T(1)) ? ga[at] : T(0);
      });
      T* ggw = gC;
      EDA_HIP(hipMemsetAsync(ggw, 0, sizeof(T) * F_, s_));
      op_vjp(md, m, 3, gw_, gA0, nullptr, ggw, on, md.rw);
      const T t0 = static_cast<T>(c.t0);
      const double dd = c.t1 - c.t0;
      const T den = static_cast<T>(dd >= 1e-12 ? dd : 1e-12);
      const T* V2 = V2_;
      const T* R = R_;
      T* gV2 = gD;
      each(s_, g, on, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        const T Uv = V2[at];
        T rho = T(0);
        const bool big = static_cast<double>(Uv) > dtol[b];
        T ratio = T(0);
        if (big) {
          ratio = R[at] / Uv;
          rho = clip_np<T>(ratio, T(0), T(1));
        }
        const T u = (rho - t0) / den;
        const T grho = (u >= T(0) && u <= T(1)) ? ggw[at] / den : T(0);
        T gr_ = T(0), gv = T(0);
        if (big && ratio >= T(0) && ratio <= T(1)) {
          gr_ = grho / Uv;
          gv = -grho * ratio / Uv;
        }
        gR[at] = gr_;
        gV2[at] = gv;
      });
      T* gV1 = gE;
      EDA_HIP(hipMemsetAsync(gV1, 0, sizeof(T) * F_, s_));
      op_vjp(md, m, 0, V1_, gV2, nullptr, gV1, on, md.rw);
      op_vjp(md, m, 0, R_, gV1, nullptr, gR, on, md.rw);
    }
    // sift 1: h1 = r - Avg(mid1), mid1 = 0.5 (u1 + l1), (u1, l1) = DilEro(r), R = max(u1 - l1, 0)
    T* gr = gr_;
    {
      T* gm1 = gB;
      each(s_, g, on, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        gr[at] = gr[at] + gh[at];
        gm1[at] = -gh[at];
      });
      T* gmid1 = gC;
      EDA_HIP(hipMemsetAsync(gmid1, 0, sizeof(T) * F_, s_));
      op_vjp(md, m, 3, mid1_, gm1, nullptr, gmid1, on, md.rw);
      const T* u1 = u1_;
      const T* l1 = l1_;
      T* gu = gD;
      T* gl = gE;
      each(s_, g, on, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        const T half = T(0.5) * gmid1[at];
        const T dv = u1[at] - l1[at];
        const T grr = dv >= T(0) ? gR[at] : T(0);
        gu[at] = half + grr;
        gl[at] = half - grr;
      });
      op_vjp(md, m, 2, r, gu, gl, gr, on, md.rw);
    }
    // the orientation: t and o through c and n, then n and c back to r
    if (orient_grad_) {
      const T* L = static_cast<const T*>(tm.L.p);
      EDA_GO(k_element_vjp<T>, elem_grid_t(g), dim3(kThreads), 0, s_, g, static_cast<T>(c.q),
             static_cast<T>(c.span - 1.0), L, coh_, nvec_, md.el, lo_, gt_, go_, gc_, gn_, on);
      EDA_LAUNCH("k_element_vjp");
      enqueue_orientation_vjp<T>(s_, dev_, g, r, md.w, on, gn_, gc_, gr, owork_, dscr_);
    }
  }

  hipStream_t s_;
  int dev_;
  MethodCtl c_;
  const ChunkTape& ct_;
  Geom g_;
  bool orient_grad_;
  int64_t B_ = 0, F_ = 0;
  int sifts_ = 0, K_max_ = 1, nsift_f_ = 0;
  TapeBuf mem_, params_;
  std::vector<ModeDev> md_;
  const uint8_t* lo_ = nullptr;
  T *nvec_, *coh_, *o_, *t_, *owork_;
  T *u1_, *l1_, *mid1_, *R_, *h1_, *V1_, *V2_, *gw_, *A0_, *a_, *m1_, *hlast_;
  std::vector<std::vector<T*>> sf_;
  T *gh_, *ghj_, *gr_, *ga_, *gt_, *go_, *gn_, *gc_;
  static constexpr int kScratch = 15;
  T* gs_[kScratch];
  T *r0_, *r1_, *gC0_, *gC1_, *Z_, *G_;
  double* dscr_;
  uint64_t* dmm_;
  int32_t* nonfin_ = nullptr;
  int32_t* kexp_ = nullptr;
  int32_t* nimfs_dev_ = nullptr;
};

}  // namespace

template <class T>
void adjoint_chunk(hipStream_t s, int device, const Tape& tape, const ChunkTape& ct, const T* grad_rows, int32_t P,
                   T* grad_x, bool orient_grad) {
  Adjoint<T> adj(s, device, tape.c, ct, orient_grad);
  adj.backward(grad_rows, P, grad_x);
  EDA_HIP(hipStreamSynchronize(s));
}
template <class T>
void replay_chunk(hipStream_t s, int device, const Tape& tape, const ChunkTape& ct, const T* x, T* rows, int32_t P) {
  Adjoint<T> adj(s, device, tape.c, ct, true);
  adj.replay(x, rows, P);
  EDA_HIP(hipStreamSynchronize(s));
}
template void adjoint_chunk<float>(hipStream_t, int, const Tape&, const ChunkTape&, const float*, int32_t, float*, bool);
template void adjoint_chunk<double>(hipStream_t, int, const Tape&, const ChunkTape&, const double*, int32_t, double*,
                                    bool);
template void replay_chunk<float>(hipStream_t, int, const Tape&, const ChunkTape&, const float*, float*, int32_t);
template void replay_chunk<double>(hipStream_t, int, const Tape&, const ChunkTape&, const double*, double*, int32_t);

}  // namespace eda
