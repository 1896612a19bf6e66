// SPDX-License-Identifier: Apache-2.0
// The separable cube filter: running max, min or box mean of an odd window along every axis, by
// van Herk / Gil-Werman passes (reference: cpp/src/field/vhgw.cu:1-36 for the canonical order,
// cpp/src/method/common.cpp:65-136 for the cube; host twin cpp/tests/twin.hpp:74-133).
//
// One line of n samples, window w: r = min((w-1)/2, n-1), wa = 2r+1; the extended line e in
// [-r, n+r) reads the line through half-sample reflection.
//
// Max and min are exact, so their result does not depend on how the line is cut: with r <= n-1
// every reflected sample of a window is also inside the clipped window [max(0, x-r), min(n-1, x+r)],
// so y(x) is the max (min) of the clipped window. Here the line is cut into blocks [jL, jL+L),
// L = wa, aligned at 0: H = the suffix folds and G = the prefix folds inside the blocks, and
// y(x) = op(H(x-r), G(x+r)) except at the clipped ends (x < r: G(x+r); a window inside the last,
// truncated block: H(x-r)). The folds compare integer order keys.
//
// The box sum keeps the reference's canonical order, which fixes its bits: segments of wa samples
// starting at o + k wa, o = (n (wa+1)/2) mod wa; the prefix fold G of a segment runs in 32-sample
// chunks (a sequential fold inside a chunk, a sequential carry over the chunk totals, G = carry +
// local); the suffix fold H is G of the reversed line; y(x) = H(x-r) + G(x+r), times 0.5 when x+r
// ends a segment (the window is the segment).
//
// gfx950 mapping. Max / min keep each lane's run of KT consecutive samples in registers, with block
// resets from a bitmask and a segmented scan across the lanes (DPP row shifts inside 16-lane rows,
// readlane across rows, only as deep as a block is long); G and H meet in LDS for the output.
// Along the contiguous axis (n <= 2048) this is a streaming kernel, one wave per line, loading the
// next line while it folds the current one (k_vhgw_rows). Otherwise (strided axes, long lines, and
// every box sum) a workgroup of W wave64s stages P lines in LDS with coalesced global accesses (P
// consecutive rows, or P adjacent columns: 16-byte loads over 64-byte runs), one pad word per 32
// samples, and writes the outputs back the same way (k_vhgw). The box sum runs one 32-sample chunk
// per lane, with the chunk tasks of all the lines a wave owns packed into its 64-lane rounds and the
// chunk carries passed lane to lane by DPP wave shifts; the G pass adds H and writes y in place.
// Workgroup ids are remapped so that consecutive groups of one element run on one XCD and share its
// L2.
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <map>
#include <mutex>
#include <atomic>
#include <limits>
#include <string>
#include <vector>

#ifndef EDA_VHGW_SUMU
#define EDA_VHGW_SUMU 4  // 16-byte reads in flight per thread while a box-sum workgroup stages its lines
#endif
#ifndef EDA_VHGW_MMU
#define EDA_VHGW_MMU 2  // the same for max / min
#endif
#ifndef EDA_VHGW_PIPE
#define EDA_VHGW_PIPE 1  // box sums through the pipelined kernel where its tile allows
#endif
#ifndef EDA_BOX_ATTR
#define EDA_BOX_ATTR  // extra attributes of the pipelined box kernel (occupancy hints)
#endif

namespace eda {
namespace {

constexpr int kVMax = 0, kVMin = 1, kVSum = 2;

// LDS slot of sample i of a staged line: one pad word per 32 samples, so that lanes 32 samples
// apart (the chunk walks), lanes KT samples apart (the register runs) and the columns of a staged
// tile fall in distinct banks.
__host__ __device__ __forceinline__ int32_t pidx(int32_t i) { return i + (i >> 5); }

// i / L and i mod L for 0 <= i < 2^24 by a float reciprocal and one correction.
__device__ __forceinline__ int32_t div_small(int32_t i, int32_t L, float inv) {
  int32_t q = static_cast<int32_t>(static_cast<float>(i) * inv);
  const int32_t m = i - q * L;
  q += m < 0 ? -1 : (m >= L ? 1 : 0);
  return q;
}
__device__ __forceinline__ int32_t mod_small(int32_t i, int32_t L, float inv) {
  int32_t m = i - static_cast<int32_t>(static_cast<float>(i) * inv) * L;
  m += m < 0 ? L : (m >= L ? -L : 0);
  return m;
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// k = first, first + stride, ... < end: U reads issued back to back before any is used (the memory
// parallelism a tile's staging needs: one read in flight per thread leaves HBM mostly idle).
template <int U, class Read, class Use>
__device__ __forceinline__ void batched(int32_t first, int32_t end, int32_t stride, Read read, Use use) {
  using Q = decltype(read(first));
  for (int32_t k0 = first; k0 < end; k0 += U * stride) {
    Q q[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (k0 + u * stride < end) q[u] = read(k0 + u * stride);
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (k0 + u * stride < end) use(k0 + u * stride, q[u]);
  }
}

// ---- cross-lane moves -----------------------------------------------------------------------------
// DPP (gfx9 encodings): row_shr:d gives lane i the value of lane i-d of its 16-lane row, row_shl:d
// that of lane i+d; wave_shr:1 / wave_shl:1 shift by one lane across the wave. Lanes whose source
// lies outside the row (wave) keep `old`.
constexpr int kRowShl = 0x100, kRowShr = 0x110, kWaveShl1 = 0x130, kWaveShr1 = 0x138;

template <int CTRL>
__device__ __forceinline__ int32_t dpp32(int32_t old, int32_t v) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, 0xf, 0xf, false);
}
template <int CTRL, class S>
__device__ __forceinline__ S dpp(S old, S v) {
  if constexpr (sizeof(S) == 4) {
    return __builtin_bit_cast(S, dpp32<CTRL>(__builtin_bit_cast(int32_t, old), __builtin_bit_cast(int32_t, v)));
  } else {
    const uint64_t o = __builtin_bit_cast(uint64_t, old), x = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = static_cast<uint32_t>(
        dpp32<CTRL>(static_cast<int32_t>(static_cast<uint32_t>(o)), static_cast<int32_t>(static_cast<uint32_t>(x))));
    const uint32_t hi = static_cast<uint32_t>(dpp32<CTRL>(static_cast<int32_t>(static_cast<uint32_t>(o >> 32)),
                                                          static_cast<int32_t>(static_cast<uint32_t>(x >> 32))));
    return __builtin_bit_cast(S, (static_cast<uint64_t>(hi) << 32) | lo);
  }
}
template <class S>
__device__ __forceinline__ S readlane(S v, int l) {
  if constexpr (sizeof(S) == 4) {
    return __builtin_bit_cast(S, __builtin_amdgcn_readlane(__builtin_bit_cast(int32_t, v), l));
  } else {
    const uint64_t x = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int32_t>(static_cast<uint32_t>(x)), l));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int32_t>(static_cast<uint32_t>(x >> 32)), l));
    return __builtin_bit_cast(S, (static_cast<uint64_t>(hi) << 32) | lo);
  }
}

template <int OP, class S>
__device__ __forceinline__ S mm_op(S a, S b) {
  if constexpr (OP == kVMax) return a > b ? a : b;
  else return a < b ? a : b;
}
template <int OP, class S>
__device__ __forceinline__ S mm_ident() {
  if constexpr (OP == kVMax) return std::numeric_limits<S>::lowest();
  else return std::numeric_limits<S>::max();
}

// One doubling step of a segmented scan: (a, f) takes the lane D back (UP) or ahead (!UP) in its
// row unless f is set.
template <bool UP, int D, int OP, class S>
__device__ __forceinline__ void seg_step(S& a, int& f, S id) {
  constexpr int ctrl = UP ? kRowShr + D : kRowShl + D;
  const S a2 = dpp<ctrl>(id, a);
  const int f2 = dpp32<ctrl>(0, f);
  if (!f) {
    a = mm_op<OP>(a2, a);
    f = f2;
  }
}
// Inclusive segmented scan over the wave: a lane's (a, f) becomes the fold of a over the lanes from
// the nearest lane with f = 1 to itself, in lane order (UP) or reverse lane order. depth: the
// doubling steps inside a row (chains of up to 2^depth - 1 lanes back); then the rows in order.
template <bool UP, int OP, class S>
__device__ __forceinline__ void seg_scan(S& a, int& f, int depth, S id) {
  if (depth > 0) seg_step<UP, 1, OP>(a, f, id);
  if (depth > 1) seg_step<UP, 2, OP>(a, f, id);
  if (depth > 2) seg_step<UP, 4, OP>(a, f, id);
  if (depth > 3) seg_step<UP, 8, OP>(a, f, id);
  const int row = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int k = 1; k < 4; ++k) {
    const int rr = UP ? k : 3 - k;                 // the row that takes its neighbour's value
    const int src = UP ? 16 * rr - 1 : 16 * rr + 16;  // the neighbour row's last (first) lane
    const S as = readlane(a, src);
    const int fs = readlane(f, src);
    if (row == rr && !f) {
      a = mm_op<OP>(as, a);
      f = fs;
    }
  }
}

// ---- max / min of one line (one wave) -----------------------------------------------------------------
// V: the staged keys, valid up to the tile end (samples beyond n repeat x[n-1]); Hb, Gb: the
// wave's rows; Y: the output slots (over V).
// The per-line constants of the block folds: block length L = 2r+1, the block starts of a run of KT
// samples as a bitmask (the multiples of L, shifted to the run's first start), and the segmented
// scan's depth: a chain of lanes without a start is at most (L - 1) / KT long.
struct MmCtl {
  int32_t n, r, L;
  float inv;
  uint32_t pat;
  int32_t chain;
  int depth;
};
template <int KT>
__device__ __forceinline__ MmCtl mm_ctl(int32_t n, int32_t r) {
  MmCtl c;
  c.n = n;
  c.r = r;
  c.L = 2 * r + 1;
  c.inv = 1.0f / static_cast<float>(c.L);
  c.pat = 0;
  for (int32_t j = 0; j < KT; j += c.L) c.pat |= 1u << j;
  c.chain = (c.L - 1) / KT;
  c.depth = c.chain == 0 ? 0 : c.chain < 2 ? 1 : c.chain < 4 ? 2 : c.chain < 8 ? 3 : 4;
  return c;
}
template <int KT>
__device__ __forceinline__ uint32_t mm_starts(const MmCtl& c, int32_t i0) {
  constexpr uint32_t kMask = KT >= 32 ? 0xffffffffu : ((1u << KT) - 1u);
  const int32_t m0 = mod_small(i0, c.L, c.inv);
  const int32_t k0 = m0 == 0 ? 0 : c.L - m0;
  return k0 < KT ? (c.pat << k0) & kMask : 0u;
}

// H (suffix folds inside the blocks) of one tile whose lane run starts at i0, the keys in v; the
// tile to the right is folded in through carry (its H at its first sample). Writes Hb[pidx(i0 + k)]
// and returns this tile's H at its first sample.
template <int OP, class S, int KT>
__device__ __forceinline__ S mm_tile_h(const S (&v0)[KT], int32_t i0, S carry, S* __restrict__ Hb, const MmCtl& c) {
  const S id = mm_ident<OP, S>();
  const uint32_t starts = mm_starts<KT>(c, i0);
  const uint32_t ends = (starts >> 1) | (mod_small(i0 + KT, c.L, c.inv) == 0 ? 1u << (KT - 1) : 0u);
  S v[KT];
  S h = id;
#pragma unroll
  for (int k = KT - 1; k >= 0; --k) {
    h = (ends >> k) & 1u ? v0[k] : mm_op<OP>(v0[k], h);
    v[k] = h;
  }
  S a = v[0];
  int f = ends != 0;
  if (c.chain > 0) seg_scan<false, OP>(a, f, c.depth, id);
  if (!f) a = mm_op<OP>(a, carry);
  const S nxt = dpp<kWaveShl1>(carry, a);
  const int kl = ends ? 31 - __clz(static_cast<int>(ends)) : -1;
  S* dst = Hb + pidx(i0);
#pragma unroll
  for (int k = 0; k < KT; ++k) dst[k] = k > kl ? mm_op<OP>(v[k], nxt) : v[k];
  return readlane(a, 0);
}
// G (prefix folds inside the blocks) of one tile; carry: the tile to the left's G at its last
// sample. Writes Gb[pidx(i0 + k)] and returns this tile's G at its last sample.
template <int OP, class S, int KT>
__device__ __forceinline__ S mm_tile_g(const S (&v0)[KT], int32_t i0, S carry, S* __restrict__ Gb, const MmCtl& c) {
  const S id = mm_ident<OP, S>();
  const uint32_t starts = mm_starts<KT>(c, i0);
  S v[KT];
  S g = id;
#pragma unroll
  for (int k = 0; k < KT; ++k) {
    g = (starts >> k) & 1u ? v0[k] : mm_op<OP>(g, v0[k]);
    v[k] = g;
  }
  S a = v[KT - 1];
  int f = starts != 0;
  if (c.chain > 0) seg_scan<true, OP>(a, f, c.depth, id);
  if (!f) a = mm_op<OP>(carry, a);
  const S prv = dpp<kWaveShr1>(carry, a);
  const int kf = starts ? __ffs(static_cast<int>(starts)) - 1 : KT;
  S* dst = Gb + pidx(i0);
#pragma unroll
  for (int k = 0; k < KT; ++k) dst[k] = k < kf ? mm_op<OP>(prv, v[k]) : v[k];
  return readlane(a, 63);
}
// y(x) for the lane's x = lane + 64 j < n from H and G: emit(x, y).
template <int OP, class S, class Emit>
__device__ __forceinline__ void mm_out(const S* __restrict__ Hb, const S* __restrict__ Gb, const MmCtl& c, Emit emit) {
  const int lane = threadIdx.x & 63;
  const int32_t n = c.n, r = c.r;
  const int32_t last = div_small(n - 1, c.L, c.inv) * c.L;  // the last block's start
  for (int32_t x0 = 0; x0 < n; x0 += 64) {
    const int32_t x = x0 + lane;
    if (x0 >= r && x0 + 63 + r <= n - 1) {  // every window of the 64 lies inside the line
      emit(x, mm_op<OP>(Hb[pidx(x - r)], Gb[pidx(x + r)]));
    } else if (x < n) {
      const int32_t lo = x - r > 0 ? x - r : 0;
      const int32_t hi = x + r < n - 1 ? x + r : n - 1;
      const S h = Hb[pidx(lo)];
      const S g = Gb[pidx(hi)];
      S y;
      if (x < r) y = g;
      else if (x + r > n - 1 && lo >= last) y = h;
      else y = mm_op<OP>(h, g);
      emit(x, y);
    }
  }
}

// mm_out into registers: y[j] = y(64 j + lane) (junk where that is n or more).
// The max (H0, G0) and the min (H1, G1) of each sample side by side, as mm_out computes each.
template <class S, class Emit2>
__device__ __forceinline__ void mm_out_pair(const S* __restrict__ H0, const S* __restrict__ G0,
                                            const S* __restrict__ H1, const S* __restrict__ G1, const MmCtl& c,
                                            Emit2 emit) {
  const int lane = threadIdx.x & 63;
  const int32_t n = c.n, r = c.r;
  const int32_t last = div_small(n - 1, c.L, c.inv) * c.L;  // the last block's start
  for (int32_t x0 = 0; x0 < n; x0 += 64) {
    const int32_t x = x0 + lane;
    if (x0 >= r && x0 + 63 + r <= n - 1) {  // every window of the 64 lies inside the line
      emit(x, mm_op<kVMax>(H0[pidx(x - r)], G0[pidx(x + r)]), mm_op<kVMin>(H1[pidx(x - r)], G1[pidx(x + r)]));
    } else if (x < n) {
      const int32_t lo = x - r > 0 ? x - r : 0;
      const int32_t hi = x + r < n - 1 ? x + r : n - 1;
      const S h0 = H0[pidx(lo)], g0 = G0[pidx(hi)];
      const S h1 = H1[pidx(lo)], g1 = G1[pidx(hi)];
      S y0, y1;
      if (x < r) {
        y0 = g0;
        y1 = g1;
      } else if (x + r > n - 1 && lo >= last) {
        y0 = h0;
        y1 = h1;
      } else {
        y0 = mm_op<kVMax>(h0, g0);
        y1 = mm_op<kVMin>(h1, g1);
      }
      emit(x, y0, y1);
    }
  }
}

template <int OP, class S, int KT>
__device__ __forceinline__ void mm_out_regs(const S* __restrict__ Hb, const S* __restrict__ Gb, const MmCtl& c,
                                            S (&y)[KT]) {
  const int lane = threadIdx.x & 63;
  const int32_t n = c.n, r = c.r;
  const int32_t last = div_small(n - 1, c.L, c.inv) * c.L;
#pragma unroll
  for (int j = 0; j < KT; ++j) {
    const int32_t x = 64 * j + lane;
    const int32_t lo = x - r > 0 ? x - r : 0;
    const int32_t hi = x + r < n - 1 ? x + r : n - 1;
    const S h = Hb[pidx(min(lo, n - 1))];
    const S g = Gb[pidx(hi)];
    if (x < r) y[j] = g;
    else if (x + r > n - 1 && lo >= last) y[j] = h;
    else y[j] = mm_op<OP>(h, g);
  }
}

// A staged line of any length (tiles of 64 KT samples): V valid up to the tile end (samples beyond
// n repeat x[n-1]); Y: the output slots (over V).
template <int OP, class S, int KT>
__device__ void mm_line(const S* __restrict__ V, S* __restrict__ Hb, S* __restrict__ Gb, S* Y, int32_t n,
                        int32_t r) {
  const int lane = threadIdx.x & 63;
  const MmCtl c = mm_ctl<KT>(n, r);
  const S id = mm_ident<OP, S>();
  constexpr int32_t kTile = 64 * KT;
  const int32_t ntiles = (n + kTile - 1) / kTile;
  S carry = id;
  for (int32_t t = ntiles - 1; t >= 0; --t) {
    const int32_t i0 = t * kTile + lane * KT;
    S v[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) v[k] = V[pidx(i0) + k];
    carry = mm_tile_h<OP, S, KT>(v, i0, carry, Hb, c);
  }
  carry = id;
  for (int32_t t = 0; t < ntiles; ++t) {
    const int32_t i0 = t * kTile + lane * KT;
    S v[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) v[k] = V[pidx(i0) + k];
    carry = mm_tile_g<OP, S, KT>(v, i0, carry, Gb, c);
  }
  wave_sync();
  mm_out<OP, S>(Hb, Gb, c, [&](int32_t x, S y) { Y[pidx(x)] = y; });
  wave_sync();
}

// ---- the canonical box sum of one line (one wave) --------------------------------------------------------
struct WinInfo {
  int32_t r, wa, o, k_lo, nsegs, cs;
};
__host__ __device__ __forceinline__ WinInfo win_info(int32_t n, int32_t w) {
  WinInfo v;
  const int32_t half = (w - 1) / 2;
  v.r = half < n - 1 ? half : n - 1;
  v.wa = 2 * v.r + 1;
  v.o = static_cast<int32_t>((static_cast<int64_t>(n) * ((v.wa + 1) / 2)) % v.wa);
  v.k_lo = static_cast<int32_t>(floor_div(static_cast<int64_t>(v.r) - v.o, v.wa));
  const int32_t k_hi = static_cast<int32_t>(floor_div(static_cast<int64_t>(n) - 1 + v.r - v.o, v.wa));
  v.nsegs = k_hi - v.k_lo + 1;
  v.cs = (v.wa + 31) / 32;
  return v;
}

// The box sums of the nl lines of a tile. Line p is staged at E + p epad + elead: the extended
// position e in [-r, n + r) at offset e, with at least 32 slots of slack past both ends; its H row
// (H + p epad + elead) has the same shape. Windows of at most 31 samples (r < 16): a segment is one
// 32-sample chunk of the canonical order, so G is the sequential fold from the segment's start. One
// lane per segment (line, segment k), the tasks of the tile spread over the workgroup's waves in
// rounds of 64; each lane walks its wa samples in order, the first wa mod 8 through a fallthrough
// chain and the rest in unrolled groups of eight (no per-sample loop control). All lanes of a round
// step together, so the words they touch lie wa (odd) apart. A segment that starts past the end of
// the extended line is skipped; one that runs past it reads and writes the slack there. DIR 1 walks
// the reversed line (its position e' reads e = n - 1 - e')
// and writes H(e') at H row offset e, so H(x - r) sits at offset x - r; DIR 0 walks the line and
// writes y(x) = H(x - r) + G(x + r) over its input at offset e = x + r, halved where e ends its
// segment (every input sample is read by the one lane whose segment holds it, in the step that
// overwrites it). Stores need no mask: positions without an output are the lane's own slack or
// samples.
template <int DIR, class T>
__device__ void box_dir(T* __restrict__ E, T* __restrict__ H, int32_t epad, int32_t elead, int32_t nl, int32_t n,
                        const WinInfo& wi) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int W = blockDim.x >> 6;
  // a segment here is at most 31 samples long (r < 16): one chunk, so G is the plain running fold
  const int32_t r = wi.r, wa = wi.wa, nsegs = wi.nsegs;
  const int32_t e_max = n - 1 + r;
  const int32_t total = nl * nsegs;
  const float inv_ns = 1.0f / static_cast<float>(nsegs);
  const int32_t rem = wa & 7;  // the steps before the groups of eight
  for (int32_t u0 = wave * 64; u0 < total; u0 += W * 64) {
    const int32_t u = u0 + lane;
    const int32_t p = min(div_small(u, nsegs, inv_ns), nl - 1);
    const int32_t k = u - p * nsegs;
    const int32_t s = wi.o + (wi.k_lo + k) * wa;  // the segment's first position
    if (u >= total || s > e_max) continue;
    // sample j of the walk: position s + j, at ep[-j] (DIR 1) or ep[j] (DIR 0)
    T* ep = E + p * epad + elead + (DIR ? n - 1 - s : s);
    T* hp = H + p * epad + elead + (DIR ? n - 1 - s : s - 2 * r);
    T acc = -T(0);
    {
      // steps j = 0 .. rem - 1 as j = rem - i from bases rem samples along: entering the chain at
      // case rem runs them in order
      T* e2 = DIR ? ep - rem : ep + rem;
      T* h2 = DIR ? hp - rem : hp + rem;
      const auto first = [&](int i) {
        acc = acc + (DIR ? e2[i] : e2[-i]);
        if constexpr (DIR == 1) {
          h2[i] = acc;
        } else {
          e2[-i] = h2[-i] + acc;
        }
      };
      switch (rem) {
        case 7: first(7); [[fallthrough]];
        case 6: first(6); [[fallthrough]];
        case 5: first(5); [[fallthrough]];
        case 4: first(4); [[fallthrough]];
        case 3: first(3); [[fallthrough]];
        case 2: first(2); [[fallthrough]];
        case 1: first(1); [[fallthrough]];
        default: break;
      }
    }
    for (int32_t j = rem; j < wa; j += 8) {
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        acc = acc + (DIR ? ep[-(j + v)] : ep[j + v]);
        if constexpr (DIR == 1) {
          hp[-(j + v)] = acc;
        } else {
          ep[j + v] = hp[j + v] + acc;
        }
      }
    }
    // DIR 0: the window that is the segment
    if (DIR == 0) ep[wa - 1] = ep[wa - 1] * T(0.5);
  }
}

// The box sums of a tile's nl lines (all the workgroup's threads); y(x) of line p ends at
// E + p epad + elead + r + x.
template <class T>
__device__ void box_lines(T* E, T* H, int32_t epad, int32_t elead, int32_t nl, int32_t n, const WinInfo& wi) {
  box_dir<1, T>(E, H, epad, elead, nl, n, wi);
  __syncthreads();
  box_dir<0, T>(E, H, epad, elead, nl, n, wi);
}

// The box sums of windows longer than 32 samples (segments of several chunks), for the lines one
// wave owns: one lane per chunk, the chunk tasks of the lines packed into 64-lane rounds and the
// chunk carries passed lane to lane by DPP wave shifts (a chain never leaves the wave). The rows are
// padded (sample e at pidx(e + lead)), so that the chunks of one segment, 32 samples apart, fall in
// distinct banks. Line li < nl of the wave is staged reflected at E + p epad, p = p0 + li pstep,
// its output row at R + p lpad. DIR 1 writes H (the G of the reversed line) at slot x = e + r of the
// forward position e; DIR 0 computes G at e = x + r and writes y = H + G over it, halved where e
// ends a segment. One chunk per lane, CH >= the longest chunk: its samples are consecutive in the
// extended row, so each one's slot is a base, the sample's index as an immediate offset, and one
// more pad word past the chunk's 32-sample boundary. Samples past a chunk's end add -0, which
// leaves every fold's bits unchanged; lanes with no output touch the junk slot.
template <int DIR, int CH, class T>
__device__ void chunk_dir(const T* __restrict__ E, int32_t epad, int32_t lead, T* __restrict__ R, int32_t lpad,
                        int32_t junk, int32_t p0, int32_t pstep, int32_t nl, int32_t n, const WinInfo& wi) {
  const int lane = threadIdx.x & 63;
  const int32_t r = wi.r, wa = wi.wa;
  const int32_t per = wi.nsegs * wi.cs;  // tasks per line
  const int32_t total = nl * per;
  const float inv_cs = 1.0f / static_cast<float>(wi.cs);
  const float inv_per = 1.0f / static_cast<float>(per);
  const int iters = wi.cs - 1 < 64 ? wi.cs - 1 : 64;
  const T nz = -T(0);
  const int32_t e_max = n - 1 + r;
  T carry = nz;
  for (int32_t t0 = 0; t0 < total; t0 += 64) {
    const int32_t u = t0 + lane;
    const int32_t li = min(div_small(u, per, inv_per), nl - 1);
    const int32_t t = u - li * per;
    const int32_t kk = div_small(t, wi.cs, inv_cs);
    const int32_t q = t - kk * wi.cs;
    const int32_t s = wi.o + (wi.k_lo + kk) * wa;
    const int32_t e_last = min(s + wa - 1, e_max);
    const int32_t e_al = s + wa - 1;  // the segment's end
    const int32_t c0 = s + 32 * q;
    const int32_t len = u < total ? min(32, e_last - c0 + 1) : 0;
    const int32_t ca = min(c0, e_max);  // the chunk's anchor (lanes without a chunk read in range)
    const int32_t p = p0 + li * pstep;
    // sample j: extended index ib + j (DIR 0) or ib - j (DIR 1)
    const int32_t ib = DIR ? n - 1 - ca + lead : ca + lead;
    const T* row0 = E + p * epad + pidx(ib);
    const T* row1 = DIR ? row0 - 1 : row0 + 1;
    const int32_t jb = DIR ? (ib & 31) : 32 - (ib & 31);
    // samples past the chunk's end fold in too: only samples with no output, and totals no
    // output uses (a segment's last chunk, chunks past the line's end), see them
    T v[CH];
    T acc = nz;
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const T* rp = (DIR ? j > jb : j >= jb) ? row1 : row0;
      acc = acc + (DIR ? rp[-j] : rp[j]);
      v[j] = acc;
    }
    // the carries: out = the fold of the segment up to this chunk's end (chunk 0: its total);
    // cin = the previous chunk's out (-0 for chunk 0: adding it changes no bits); lane 0 continues
    // the previous round's lane 63
    T out = acc, cin = nz;
    for (int it = 0; it < iters; ++it) {
      const T prev = dpp<kWaveShr1>(carry, out);
      if (q > 0) {
        cin = prev;
        out = prev + acc;
      }
    }
    carry = readlane(out, 63);
    // output slot x of sample j: xb - j (DIR 1) or xb + j (DIR 0), valid when the sample is in its
    // chunk and e = c0 + j >= r
    const int32_t xb = DIR ? n - 1 + r - ca : ca - r;
    T* out0 = R + p * lpad + xb + (xb >> 5);
    T* out1 = DIR ? out0 - 1 : out0 + 1;
    const int32_t xm = DIR ? (xb & 31) : 32 - (xb & 31);
    const int32_t jmin = r - c0;  // the first j with an output
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const bool valid = j < len && j >= jmin;
      T* slot = (DIR ? j > xm : j >= xm) ? out1 : out0;
      T* at = valid ? (DIR ? slot - j : slot + j) : R + junk;
      const T gv = cin + v[j];
      if constexpr (DIR == 1) {
        *at = gv;
      } else {
        *at = *at + gv;
      }
    }
    if (DIR == 0) {
      // the window that is the segment: halve its sum
      const int32_t ja = e_al - c0;
      if (ja < len && ja >= 0 && ja >= jmin) {
        T* at = (ja >= xm ? out1 : out0) + ja;
        *at = *at * T(0.5);
      }
    }
  }
}


template <class T>
__device__ void chunk_lines(const T* E, int32_t epad, int32_t lead, T* R, int32_t lpad, int32_t junk, int32_t p0,
                            int32_t pstep, int32_t nl, int32_t n, const WinInfo& wi) {
  chunk_dir<1, 32, T>(E, epad, lead, R, lpad, junk, p0, pstep, nl, n, wi);
  wave_sync();
  chunk_dir<0, 32, T>(E, epad, lead, R, lpad, junk, p0, pstep, nl, n, wi);
  wave_sync();
}

// ---- the pass kernel ----------------------------------------------------------------------------------------
struct VPass {
  Geom g;
  int32_t n;        // line length
  int64_t step;     // sample stride along the line
  int64_t inner;    // adjacent lines (1 on the contiguous axis)
  int64_t outer;    // rows (contiguous axis) or column blocks per element
  int32_t P;        // lines per workgroup (a power of two on a strided axis)
  int32_t lpad;     // LDS slots per line
  int32_t span;     // samples per line the max / min tiles cover (n rounded up to the tile)
  int32_t flat;     // rows with pidx(p n + i) = p lpad + pidx(i)
  int64_t groups;   // workgroups per element
  int32_t nf;       // output fields: 1, or 2 (ops[0] into out[0], ops[1] into out[1])
  int32_t ops[2];
  int32_t scale;    // multiply by fl_T(1 / prod of the clamped sides) (the last box pass)
  int32_t xcd;      // remap workgroup ids: chunks of xcd consecutive groups rotate over the 8 XCDs (0: none)
  int32_t vec;      // strided axis: 16-byte global accesses over 16 / sizeof(T) adjacent columns
  int32_t wrows;    // LDS rows per wave: 1 (max / min of one tile: H in place), 2 (H and G)
  int32_t epad;     // box sum: LDS slots per reflected line (extended position e at e + lead)
  int32_t lead;     // 32 + the largest radius of the launch
  int32_t rmin;     // elements whose radius is below rmin are skipped (the streaming pass takes them)
};

// gridDim.y passes (rungs) in one launch: rung j reads win + j B, in + j in_stride and writes
// out + j out_stride.
template <class T>
struct VPtrs {
  const T* in[2];
  T* out[2];
  const int32_t* win;
  int64_t in_stride, out_stride;
};

// S: the LDS type (T for the box sum; the integer order key of T for max / min). KT: samples per
// lane of a max / min tile.
template <class T, class S, int KT>
__global__ void __launch_bounds__(512) k_vhgw(VPass a, VPtrs<T> ptr) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int64_t wg = a.xcd ? xcd_item(blockIdx.x, a.xcd) : static_cast<int64_t>(blockIdx.x);
  const int64_t b = wg / a.groups;
  const int64_t grp = wg - b * a.groups;
  const int64_t rung = blockIdx.y;
  const int32_t w = ptr.win[rung * a.g.batch + b];
  if (w < 1) return;  // uniform: the element does not run
  const int32_t n = a.n;
  const int nthr = blockDim.x;
  const int W = nthr >> 6;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  // Max / min: V = P staged lines of lpad slots; per wave wrows rows of lpad + 1 slots, H then G
  // (one row: G only, the line's H goes over its staged samples); the outputs over V. Box sum: P
  // reflected lines of epad slots, then their P H rows R of as many; the outputs over V.
  constexpr bool kSum = std::is_same_v<S, T>;
  S* V = reinterpret_cast<S*>(smem);
  S* Hb = V + a.P * a.lpad + wave * a.wrows * (a.lpad + 1);
  S* Gb = a.wrows == 2 ? Hb + a.lpad + 1 : Hb;
  S* R = V + a.P * a.epad;
  const bool contiguous = a.inner == 1;
  int64_t base0;
  int32_t lines;
  if (contiguous) {
    const int64_t l0 = grp * a.P;
    lines = static_cast<int32_t>(min<int64_t>(a.P, a.outer - l0));
    base0 = b * a.g.count + l0 * n;
  } else {
    const int64_t gpo = (a.inner + a.P - 1) / a.P;
    const int64_t o = grp / gpo;
    const int64_t j0 = (grp - o * gpo) * a.P;
    lines = static_cast<int32_t>(min<int64_t>(a.P, a.inner - j0));
    base0 = b * a.g.count + o * static_cast<int64_t>(n) * a.inner + j0;
  }
  const int32_t total = lines * n;
  const float inv_n = 1.0f / static_cast<float>(n);
  const int pshift = __ffs(a.P) - 1;  // strided axis: P is a power of two
  T fac = T(1);
  if (a.scale) {
    double prod = 1.0;
    for (int ax = 0; ax < a.g.dim; ++ax) {
      const int32_t h = (w - 1) / 2;
      prod *= static_cast<double>(2 * min(h, a.g.n[ax] - 1) + 1);
    }
    fac = static_cast<T>(1.0 / prod);
  }
  const int32_t half = (w - 1) / 2;
  const int32_t r = half < n - 1 ? half : n - 1;
  if (kSum && r > a.lead - 32) return;  // a window above the launch's largest (the host prevents it)
  if (r < a.rmin) return;               // uniform: the streaming pass of this axis serves the element
  // box sum: windows over 32 samples fold chunk-parallel on padded rows, others walk their segments
  // on plain rows; the offset of extended position e in a staged row
  const bool padded = kSum && r >= 16;
  const auto eslot = [&](int32_t e) { return padded ? pidx(e + a.lead) : e + a.lead; };
  constexpr int VW = 16 / static_cast<int>(sizeof(T));
  using V16 = std::conditional_t<sizeof(T) == 4, float4, double2>;
  constexpr int kU = kSum ? EDA_VHGW_SUMU : EDA_VHGW_MMU;  // staging reads in flight per thread
  constexpr int kO = kSum ? EDA_VHGW_SUMU : EDA_VHGW_MMU;  // write-out rows per thread between waits
  for (int f = 0; f < a.nf; ++f) {
    const T* src = ptr.in[f] + rung * ptr.in_stride;
    // sample i of line p (the box sum: at its extended position e = i)
    const auto put = [&](int32_t p, int32_t i, T v) {
      if constexpr (kSum) {
        V[p * a.epad + eslot(i)] = v;
      } else {
        V[p * a.lpad + pidx(i)] = okey(v);
      }
    };
    if (a.flat) {
      // max / min rows in one run: pidx(p n + i) = p lpad + pidx(i)
      if constexpr (!kSum)
        batched<kU>(
            threadIdx.x, total, nthr, [&](int32_t idx) { return src[base0 + idx]; },
            [&](int32_t idx, T v) { V[pidx(idx)] = okey(v); });
    } else if (contiguous && a.vec) {
      // 16-byte loads of VW consecutive samples of one line (n is a multiple of VW)
      batched<kU>(
          threadIdx.x * VW, total, nthr * VW,
          [&](int32_t idx) { return *reinterpret_cast<const V16*>(src + base0 + idx); },
          [&](int32_t idx, const V16& q) {
            const int32_t p = div_small(idx, n, inv_n);
            const int32_t i = idx - p * n;
            if constexpr (VW == 4) {
              put(p, i, q.x);
              put(p, i + 1, q.y);
              put(p, i + 2, q.z);
              put(p, i + 3, q.w);
            } else {
              put(p, i, q.x);
              put(p, i + 1, q.y);
            }
          });
    } else if (contiguous) {
      batched<kU>(
          threadIdx.x, total, nthr, [&](int32_t idx) { return src[base0 + idx]; },
          [&](int32_t idx, T v) {
            const int32_t p = div_small(idx, n, inv_n);
            put(p, idx - p * n, v);
          });
    } else if (a.vec && lines == a.P) {
      // 16-byte loads: tpr threads per row, each VW adjacent columns
      const int tpr = a.P / VW;
      const int cg = threadIdx.x % tpr;
      batched<kU>(
          static_cast<int32_t>(threadIdx.x) / tpr, n, nthr / tpr,
          [&](int32_t i) {
            return *reinterpret_cast<const V16*>(src + base0 + static_cast<int64_t>(i) * a.step + cg * VW);
          },
          [&](int32_t i, const V16& q) {
            if constexpr (VW == 4) {
              put(cg * 4 + 0, i, q.x);
              put(cg * 4 + 1, i, q.y);
              put(cg * 4 + 2, i, q.z);
              put(cg * 4 + 3, i, q.w);
            } else {
              put(cg * 2 + 0, i, q.x);
              put(cg * 2 + 1, i, q.y);
            }
          });
    } else {
      const int32_t p = threadIdx.x & (a.P - 1);
      if (p < lines)
        batched<kU>(
            static_cast<int32_t>(threadIdx.x >> pshift), n, nthr >> pshift,
            [&](int32_t i) { return src[base0 + static_cast<int64_t>(i) * a.step + p]; },
            [&](int32_t i, T v) { put(p, i, v); });
    }
    __syncthreads();
    if constexpr (kSum) {
      // the reflections the window reads: e = -1 - i for i < r, e = 2n - 1 - i for i >= n - r
      if (r > 0) {
        const int32_t nref = 2 * r;
        const float inv_ref = 1.0f / static_cast<float>(nref);
        for (int32_t idx = threadIdx.x; idx < lines * nref; idx += nthr) {
          const int32_t p = div_small(idx, nref, inv_ref);
          const int32_t t = idx - p * nref;
          S* El = V + p * a.epad;
          const int32_t i = t < r ? t : n - 1 - (t - r);
          const int32_t e = t < r ? -1 - i : 2 * n - 1 - i;
          El[eslot(e)] = El[eslot(i)];
        }
        __syncthreads();
      }
      const WinInfo wi = win_info(n, w);
      if (padded) {
        const int32_t nl = wave < lines ? (lines - wave + W - 1) / W : 0;  // wave w owns lines w, w + W, ...
        chunk_lines<T>(V, a.epad, a.lead, R, a.lpad, a.P * a.lpad, wave, W, nl, n, wi);
      } else {
        box_lines<T>(V, R, a.epad, a.lead, lines, n, wi);
      }
    } else {
      for (int p = wave; p < lines; p += W) {
        S* Vl = V + p * a.lpad;
        // the tiles read up to the span: repeat x[n-1] there
        const S tail = Vl[pidx(n - 1)];
        for (int32_t i = n + lane; i < a.span; i += 64) Vl[pidx(i)] = tail;
        wave_sync();
        if (a.wrows == 1) {
          // one tile: the line in registers, H over the staged samples, G in the wave's row, the
          // outputs in registers until every lane has read H and G
          const int32_t i0 = lane * KT;
          S v[KT];
#pragma unroll
          for (int k = 0; k < KT; ++k) v[k] = Vl[pidx(i0) + k];
          const MmCtl c = mm_ctl<KT>(n, r);
          S y[KT];
          if (a.ops[f] == kVMax) {
            mm_tile_h<kVMax, S, KT>(v, i0, mm_ident<kVMax, S>(), Vl, c);
            mm_tile_g<kVMax, S, KT>(v, i0, mm_ident<kVMax, S>(), Gb, c);
            wave_sync();
            mm_out_regs<kVMax, S, KT>(Vl, Gb, c, y);
          } else {
            mm_tile_h<kVMin, S, KT>(v, i0, mm_ident<kVMin, S>(), Vl, c);
            mm_tile_g<kVMin, S, KT>(v, i0, mm_ident<kVMin, S>(), Gb, c);
            wave_sync();
            mm_out_regs<kVMin, S, KT>(Vl, Gb, c, y);
          }
          wave_sync();
#pragma unroll
          for (int j = 0; j < KT; ++j)
            if (64 * j + lane < n) Vl[pidx(64 * j + lane)] = y[j];
          wave_sync();
        } else if (a.ops[f] == kVMax) {
          mm_line<kVMax, S, KT>(Vl, Hb, Gb, Vl, n, r);
        } else {
          mm_line<kVMin, S, KT>(Vl, Hb, Gb, Vl, n, r);
        }
      }
    }
    __syncthreads();
    T* dst = ptr.out[f] + rung * ptr.out_stride;
    const auto get = [&](int32_t slot) {
      T y;
      if constexpr (kSum) y = padded ? R[slot] : V[slot];
      else y = okey_inv(V[slot]);
      return a.scale ? y * fac : y;
    };
    // the slot of output i of line p
    const auto slot = [&](int32_t p, int32_t i) {
      return kSum && !padded ? p * a.epad + a.lead + r + i : p * a.lpad + pidx(i);
    };
    if (a.flat) {
      batched<kO>(
          threadIdx.x, total, nthr, [&](int32_t idx) { return get(pidx(idx)); },
          [&](int32_t idx, T y) { dst[base0 + idx] = y; });
    } else if (contiguous && a.vec) {
      batched<kO>(
          threadIdx.x * VW, total, nthr * VW,
          [&](int32_t idx) {
            const int32_t p = div_small(idx, n, inv_n);
            const int32_t sl = slot(p, idx - p * n);  // VW samples in one 32-block: consecutive slots
            V16 q;
            if constexpr (VW == 4) {
              q.x = get(sl);
              q.y = get(sl + 1);
              q.z = get(sl + 2);
              q.w = get(sl + 3);
            } else {
              q.x = get(sl);
              q.y = get(sl + 1);
            }
            return q;
          },
          [&](int32_t idx, const V16& q) { *reinterpret_cast<V16*>(dst + base0 + idx) = q; });
    } else if (contiguous) {
      batched<kO>(
          threadIdx.x, total, nthr,
          [&](int32_t idx) {
            const int32_t p = div_small(idx, n, inv_n);
            return get(slot(p, idx - p * n));
          },
          [&](int32_t idx, T y) { dst[base0 + idx] = y; });
    } else if (a.vec && lines == a.P) {
      const int tpr = a.P / VW;
      const int cg = threadIdx.x % tpr;
      batched<kO>(
          static_cast<int32_t>(threadIdx.x) / tpr, n, nthr / tpr,
          [&](int32_t i) {
            V16 q;
            if constexpr (VW == 4) {
              q.x = get(slot(cg * 4 + 0, i));
              q.y = get(slot(cg * 4 + 1, i));
              q.z = get(slot(cg * 4 + 2, i));
              q.w = get(slot(cg * 4 + 3, i));
            } else {
              q.x = get(slot(cg * 2 + 0, i));
              q.y = get(slot(cg * 2 + 1, i));
            }
            return q;
          },
          [&](int32_t i, const V16& q) {
            *reinterpret_cast<V16*>(dst + base0 + static_cast<int64_t>(i) * a.step + cg * VW) = q;
          });
    } else {
      const int32_t p = threadIdx.x & (a.P - 1);
      if (p < lines)
        batched<kO>(
            static_cast<int32_t>(threadIdx.x >> pshift), n, nthr >> pshift,
            [&](int32_t i) { return get(slot(p, i)); },
            [&](int32_t i, T y) { dst[base0 + static_cast<int64_t>(i) * a.step + p] = y; });
    }
    __syncthreads();
  }
}

// ---- the box sum, persistent and pipelined ------------------------------------------------------------
// The same tiles as k_vhgw's box sum (P lines staged reflected), for tiles that 16-byte accesses
// cover, one launch per row layout (PAD: windows over 32 samples; each launch skips the other's
// items): each workgroup walks its share of the (rung, element, tile) items and reads the next
// running item's samples into registers (PF 16-byte reads per thread) before it sums the current
// one, so that the HBM reads of one tile overlap the LDS work of the one before. At each step of
// the walk the workgroups take the next gridDim.x items in the XCD rotation (xcd_item): chunks of
// neighbouring tiles share an XCD's L2, and every XCD works on every rung and element.
template <class T, int PF, bool PAD>
__global__ void __launch_bounds__(256) EDA_BOX_ATTR k_box(VPass a, VPtrs<T> ptr, int32_t rungs) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int VW = 16 / static_cast<int>(sizeof(T));
  using V16 = std::conditional_t<sizeof(T) == 4, float4, double2>;
  T* E = reinterpret_cast<T*>(smem);
  T* R = E + a.P * a.epad;
  const int nthr = blockDim.x;
  const int W = nthr >> 6;
  const int wave = threadIdx.x >> 6;
  const int32_t n = a.n;
  const bool contiguous = a.inner == 1;
  const float inv_n = 1.0f / static_cast<float>(n);
  const int64_t per_rung = a.g.batch * a.groups;
  const int64_t items = per_rung * rungs;
  // grid-stride over the items from the workgroup's place in the XCD rotation
  const int64_t hi = items;
  const int64_t nloc = gridDim.x;
  const int64_t vid = a.xcd ? xcd_item(blockIdx.x, a.xcd) : static_cast<int64_t>(blockIdx.x);
  // strided axis: tpr threads per row, each VW adjacent columns; rows t / tpr + u nthr / tpr
  const int tpr = a.P / VW;
  const int cg = threadIdx.x % tpr;
  const int32_t row0 = static_cast<int32_t>(threadIdx.x) / tpr;
  const int32_t rstep = nthr / tpr;

  struct Item {
    int64_t base0;  // sample 0 of line 0 in the field
    int32_t lines, w;
    int64_t rung;
  };
  const auto decode = [&](int64_t it) {
    Item d;
    d.rung = it / per_rung;
    const int64_t rem = it - d.rung * per_rung;
    const int64_t b = rem / a.groups;
    const int64_t grp = rem - b * a.groups;
    d.w = ptr.win[d.rung * a.g.batch + b];
    if (contiguous) {
      const int64_t l0 = grp * a.P;
      d.lines = static_cast<int32_t>(min<int64_t>(a.P, a.outer - l0));
      d.base0 = b * a.g.count + l0 * n;
    } else {
      const int64_t gpo = a.inner / a.P;  // the host keeps inner a multiple of P
      const int64_t o = grp / gpo;
      d.lines = a.P;
      d.base0 = b * a.g.count + o * static_cast<int64_t>(n) * a.inner + (grp - o * gpo) * a.P;
    }
    return d;
  };
  // the first running item of this kernel's class among it, it + nloc, ... (hi if none); a window
  // above the launch's largest (the host prevents it) does not run. The lanes of a wave test 64
  // items at once, so that a launch in which few elements run skips the others in few reads.
  const int lane = threadIdx.x & 63;
  const auto next_running = [&](int64_t it) {
    for (; it < hi; it += 64 * nloc) {
      const int64_t c = it + lane * nloc;
      bool run = false;
      if (c < hi) {
        const int64_t rung = c / per_rung;
        const int64_t b = (c - rung * per_rung) / a.groups;
        const int32_t w = ptr.win[rung * a.g.batch + b];
        const int32_t rw = min((w - 1) / 2, n - 1);
        run = w >= 1 && rw <= a.lead - 32 && (rw >= 16) == PAD && rw >= a.rmin;
      }
      const uint64_t m = __ballot(run);
      if (m != 0) return it + static_cast<int64_t>(__ffsll(static_cast<unsigned long long>(m)) - 1) * nloc;
    }
    return hi;
  };
  V16 buf[PF];
  const auto fetch = [&](const Item& d) {
    const T* src = ptr.in[0] + d.rung * ptr.in_stride + d.base0;
    const int32_t total = d.lines * n;
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      if (contiguous) {
        const int32_t idx = (static_cast<int32_t>(threadIdx.x) + u * nthr) * VW;
        if (idx < total) buf[u] = *reinterpret_cast<const V16*>(src + idx);
      } else {
        const int32_t i = row0 + u * rstep;
        if (i < n) buf[u] = *reinterpret_cast<const V16*>(src + static_cast<int64_t>(i) * a.step + cg * VW);
      }
    }
  };

  int64_t it = next_running(vid);
  if (it >= hi) return;
  Item cur = decode(it);
  fetch(cur);
  while (true) {
    const int32_t half = (cur.w - 1) / 2;
    const int32_t r = half < n - 1 ? half : n - 1;
    // windows over 32 samples (PAD) fold chunk-parallel on padded rows, others walk their segments
    // on plain rows; the offset of extended position e in a staged row
    constexpr bool padded = PAD;
    const auto eslot = [&](int32_t e) { return padded ? pidx(e + a.lead) : e + a.lead; };
    // stage: sample i of line p at its extended position e = i
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      if (contiguous) {
        const int32_t idx = (static_cast<int32_t>(threadIdx.x) + u * nthr) * VW;
        if (idx >= cur.lines * n) continue;
        const int32_t p = div_small(idx, n, inv_n);
        const int32_t i = idx - p * n;
        T* El = E + p * a.epad;
        if constexpr (VW == 4) {
          El[eslot(i)] = buf[u].x;
          El[eslot(i + 1)] = buf[u].y;
          El[eslot(i + 2)] = buf[u].z;
          El[eslot(i + 3)] = buf[u].w;
        } else {
          El[eslot(i)] = buf[u].x;
          El[eslot(i + 1)] = buf[u].y;
        }
      } else {
        const int32_t i = row0 + u * rstep;
        if (i >= n) continue;
        T* dst = E + cg * VW * a.epad + eslot(i);
        if constexpr (VW == 4) {
          dst[0] = buf[u].x;
          dst[a.epad] = buf[u].y;
          dst[2 * a.epad] = buf[u].z;
          dst[3 * a.epad] = buf[u].w;
        } else {
          dst[0] = buf[u].x;
          dst[a.epad] = buf[u].y;
        }
      }
    }
    __syncthreads();
    // the reflections the window reads: e = -1 - i for i < r, e = 2n - 1 - i for i >= n - r
    if (r > 0) {
      const int32_t nref = 2 * r;
      const float inv_ref = 1.0f / static_cast<float>(nref);
      for (int32_t idx = threadIdx.x; idx < cur.lines * nref; idx += nthr) {
        const int32_t p = div_small(idx, nref, inv_ref);
        const int32_t t = idx - p * nref;
        T* El = E + p * a.epad;
        const int32_t i = t < r ? t : n - 1 - (t - r);
        const int32_t e = t < r ? -1 - i : 2 * n - 1 - i;
        El[eslot(e)] = El[eslot(i)];
      }
      __syncthreads();
    }
    // the next item's reads go out now and land while this one is summed
    const int64_t nx = next_running(it + nloc);
    Item nxt{};
    if (nx < hi) {
      nxt = decode(nx);
      fetch(nxt);
    }
    {
      const WinInfo wi = win_info(n, cur.w);
      if constexpr (PAD) {
        const int32_t nl = wave < cur.lines ? (cur.lines - wave + W - 1) / W : 0;  // wave w owns lines w, w + W, ...
        chunk_lines<T>(E, a.epad, a.lead, R, a.lpad, a.P * a.lpad, wave, W, nl, n, wi);
      } else {
        box_lines<T>(E, R, a.epad, a.lead, cur.lines, n, wi);
      }
    }
    __syncthreads();
    T fac = T(1);
    if (a.scale) {
      double prod = 1.0;
      for (int ax = 0; ax < a.g.dim; ++ax) prod *= static_cast<double>(2 * min(half, a.g.n[ax] - 1) + 1);
      fac = static_cast<T>(1.0 / prod);
    }
    T* dst = ptr.out[0] + cur.rung * ptr.out_stride + cur.base0;
    // output x of line p: padded R + p lpad + pidx(x), else E + p epad + lead + r + x
    const T* Y = padded ? R : E;
    const int32_t ypad = padded ? a.lpad : a.epad;
    const auto yslot = [&](int32_t x) { return padded ? pidx(x) : a.lead + r + x; };
    const auto get = [&](int32_t slot) { return a.scale ? Y[slot] * fac : Y[slot]; };
    if (contiguous) {
      for (int32_t idx = threadIdx.x * VW; idx < cur.lines * n; idx += nthr * VW) {
        const int32_t p = div_small(idx, n, inv_n);
        const int32_t sl = p * ypad + yslot(idx - p * n);  // VW samples in one 32-block: consecutive slots
        V16 q;
        if constexpr (VW == 4) {
          q.x = get(sl);
          q.y = get(sl + 1);
          q.z = get(sl + 2);
          q.w = get(sl + 3);
        } else {
          q.x = get(sl);
          q.y = get(sl + 1);
        }
        *reinterpret_cast<V16*>(dst + idx) = q;
      }
    } else {
      for (int32_t i = row0; i < n; i += rstep) {
        const int32_t sl = cg * VW * ypad + yslot(i);
        V16 q;
        if constexpr (VW == 4) {
          q.x = get(sl);
          q.y = get(sl + ypad);
          q.z = get(sl + 2 * ypad);
          q.w = get(sl + 3 * ypad);
        } else {
          q.x = get(sl);
          q.y = get(sl + ypad);
        }
        *reinterpret_cast<V16*>(dst + static_cast<int64_t>(i) * a.step + cg * VW) = q;
      }
    }
    if (nx >= hi) break;
    __syncthreads();  // the outputs sit in the rows the next item is staged into
    it = nx;
    cur = nxt;
  }
}

// ---- the streaming pass of the contiguous axis (lines of at most 64 * 32 samples) ------------------------
// One wave per line, grid-stride over the (rung, element, line) items; each lane holds KT consecutive
// samples of the line in registers, loaded with 16-byte loads where the line allows, and the next
// item's samples are loaded while the current one is folded. Max / min only: the registers fold
// directly (G and H through the wave's two LDS rows), and the outputs go straight to global memory,
// lanes on consecutive samples.
struct RPass {
  Geom g;
  int32_t n;       // line length
  int64_t lines;   // lines per element
  int64_t items;   // rungs * batch * lines
  int32_t nf;      // output fields: 1, or 2 (ops[0] into out[0], ops[1] into out[1])
  int32_t two_in;  // nf == 2 with two inputs (else field 1 folds the input of field 0)
  int32_t ops[2];
  int32_t vec;     // n % 4 == 0: 16-byte loads
  int32_t il;      // nf == 2 (max, min): both outputs of sample v side by side at out[0] + 2 v
  int32_t rmin;    // elements whose radius is below rmin are skipped (the fused cube takes them)
};

template <class T, int KT>
struct LineRegs {
  T v[2][KT];
  T tail[2];  // x[n-1]
};

template <class T, class S, int KT>
__global__ void __launch_bounds__(256) k_vhgw_rows(RPass a, VPtrs<T> ptr) {
  static_assert(!std::is_same_v<S, T>, "max / min only");
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int32_t kSpan = 64 * KT;
  constexpr int32_t kSlots = kSpan + kSpan / 32;  // pidx(kSpan - 1) + 1
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int W = blockDim.x >> 6;
  S* Hb = reinterpret_cast<S*>(smem) + wave * (a.il ? 4 : 2) * kSlots;
  S* Gb = Hb + kSlots;
  const int32_t n = a.n;
  const int64_t B = a.g.batch;
  const int32_t i0 = lane * KT;
  const auto window = [&](int64_t it) { return ptr.win[it / a.lines]; };
  const auto load_field = [&](const T* src, T(&v)[KT], T& tail) {
    tail = src[n - 1];
    if constexpr (KT >= 4) {
      if (a.vec) {
        using V4 = std::conditional_t<sizeof(T) == 4, float4, double4>;
#pragma unroll
        for (int c = 0; c < KT / 4; ++c) {
          V4 q{};
          if (i0 + 4 * c < n) q = *reinterpret_cast<const V4*>(src + i0 + 4 * c);
          v[4 * c] = q.x;
          v[4 * c + 1] = q.y;
          v[4 * c + 2] = q.z;
          v[4 * c + 3] = q.w;
        }
        return;
      }
    }
#pragma unroll
    for (int k = 0; k < KT; ++k) v[k] = i0 + k < n ? src[i0 + k] : T(0);
  };
  const auto load = [&](int64_t it, LineRegs<T, KT>& R) {
    const int64_t rb = it / a.lines;
    const int64_t line = it - rb * a.lines;
    const int64_t rung = rb / B;
    const int64_t b = rb - rung * B;
    const int64_t off = rung * ptr.in_stride + b * a.g.count + line * n;
    load_field(ptr.in[0] + off, R.v[0], R.tail[0]);
    if (a.two_in) load_field(ptr.in[1] + off, R.v[1], R.tail[1]);
  };
  const int64_t nw = static_cast<int64_t>(gridDim.x) * W;
  // The wave's items are sb0, sb0 + nw, ...; the lanes read the windows of 64 of them at once
  // (lane l: item sb + l nw), and the running ones are taken in order from the ballot's mask, so
  // that the items of elements that do not run cost no reads of their own.
  int64_t sb = static_cast<int64_t>(blockIdx.x) * W + wave - 64 * nw;
  int32_t sw = 0;
  uint64_t mask = 0;
  const auto next = [&](int64_t& item, int32_t& w) {
    while (mask == 0) {
      sb += 64 * nw;
      if (sb >= a.items) {
        item = a.items;
        w = 0;
        return;
      }
      const int64_t c = sb + lane * nw;
      sw = c < a.items ? window(c) : 0;
      mask = __ballot(sw >= 1 && min((sw - 1) / 2, n - 1) >= a.rmin);
    }
    const int l = __ffsll(static_cast<unsigned long long>(mask)) - 1;
    mask &= mask - 1;
    item = sb + l * nw;
    w = __builtin_amdgcn_readlane(sw, l);
  };
  int64_t it = 0;
  int32_t w_cur = 0;
  next(it, w_cur);
  LineRegs<T, KT> cur{};
  if (it < a.items) load(it, cur);
  while (it < a.items) {
    int64_t it2 = 0;
    int32_t w_nxt = 0;
    next(it2, w_nxt);
    LineRegs<T, KT> nxt{};
    if (it2 < a.items) load(it2, nxt);
    {
      const int64_t rb = it / a.lines;
      const int64_t line = it - rb * a.lines;
      const int64_t rung = rb / B;
      const int64_t b = rb - rung * B;
      const int32_t half = (w_cur - 1) / 2;
      const int32_t r = half < n - 1 ? half : n - 1;
      if (a.il) {
        // the max of field 0 and the min of field 1, written side by side
        S* H1 = Gb + kSlots;
        S* G1 = H1 + kSlots;
        const MmCtl c = mm_ctl<KT>(n, r);
        S v0[KT], v1[KT];
#pragma unroll
        for (int k = 0; k < KT; ++k) {
          v0[k] = okey(i0 + k < n ? cur.v[0][k] : cur.tail[0]);
          v1[k] = okey(i0 + k < n ? (a.two_in ? cur.v[1][k] : cur.v[0][k]) : (a.two_in ? cur.tail[1] : cur.tail[0]));
        }
        mm_tile_h<kVMax, S, KT>(v0, i0, mm_ident<kVMax, S>(), Hb, c);
        mm_tile_g<kVMax, S, KT>(v0, i0, mm_ident<kVMax, S>(), Gb, c);
        mm_tile_h<kVMin, S, KT>(v1, i0, mm_ident<kVMin, S>(), H1, c);
        mm_tile_g<kVMin, S, KT>(v1, i0, mm_ident<kVMin, S>(), G1, c);
        wave_sync();
        struct alignas(2 * sizeof(T)) Pair {
          T a, b;
        };
        Pair* dst = reinterpret_cast<Pair*>(ptr.out[0] + rung * ptr.out_stride) + b * a.g.count + line * n;
        mm_out_pair<S>(Hb, Gb, H1, G1, c, [&](int32_t xo, S y0, S y1) {
          Pair q;
          q.a = okey_inv(y0);
          q.b = okey_inv(y1);
          dst[xo] = q;
        });
        wave_sync();
      }
      for (int f = 0; f < (a.il ? 0 : a.nf); ++f) {
        T* dst = ptr.out[f] + rung * ptr.out_stride + b * a.g.count + line * n;
        // the field's keys, selected without indexing the register arrays at run time; samples past
        // the line's end repeat x[n-1]
        const bool second = a.two_in && f == 1;
        const T tail = second ? cur.tail[1] : cur.tail[0];
        S v[KT];
#pragma unroll
        for (int k = 0; k < KT; ++k) v[k] = okey(i0 + k < n ? (second ? cur.v[1][k] : cur.v[0][k]) : tail);
        const MmCtl c = mm_ctl<KT>(n, r);
        const auto emit = [&](int32_t xo, S y) { dst[xo] = okey_inv(y); };
        if (a.ops[f] == kVMax) {
          mm_tile_h<kVMax, S, KT>(v, i0, mm_ident<kVMax, S>(), Hb, c);
          mm_tile_g<kVMax, S, KT>(v, i0, mm_ident<kVMax, S>(), Gb, c);
          wave_sync();
          mm_out<kVMax, S>(Hb, Gb, c, emit);
        } else {
          mm_tile_h<kVMin, S, KT>(v, i0, mm_ident<kVMin, S>(), Hb, c);
          mm_tile_g<kVMin, S, KT>(v, i0, mm_ident<kVMin, S>(), Gb, c);
          wave_sync();
          mm_out<kVMin, S>(Hb, Gb, c, emit);
        }
        wave_sync();
      }
    }
    cur = nxt;
    w_cur = w_nxt;
    it = it2;
  }
}

#ifndef EDA_VHGW_ROWS
#define EDA_VHGW_ROWS 16  // rows per workgroup on the contiguous axis
#endif
#ifndef EDA_VHGW_SUMW
#define EDA_VHGW_SUMW 4  // waves per workgroup of the box sum
#endif
#ifndef EDA_VHGW_SUMLDS
#define EDA_VHGW_SUMLDS 40000  // the LDS a box-sum workgroup aims at (bytes)
#endif
#ifndef EDA_VHGW_LDS
#define EDA_VHGW_LDS 65536  // the LDS a workgroup aims at (bytes)
#endif
#ifndef EDA_BOX_PERCU
#define EDA_BOX_PERCU 4  // resident workgroups per CU that the persistent box sum launches
#endif

int kt_for(int32_t n) {
  const int32_t per = (n + 63) / 64;
  return per <= 1 ? 1 : per <= 2 ? 2 : per <= 4 ? 4 : per <= 8 ? 8 : 16;
}

// The launch shape of one pass: P lines, W waves, LDS slots per line, the tile span.
struct Shape {
  int32_t P, W, lpad, span, flat, wrows, epad, lead;
  size_t lds;
};
// Box sums: layout 1 = plain rows only (windows up to 32 samples), 2 = padded rows only (longer
// windows), 0 = room for either (the kernel that serves both).
Shape pass_shape(const VPass& a, size_t s, bool sum, int32_t max_w, int layout = 0) {
  const int32_t n = a.n;
  const bool contiguous = a.inner == 1;
  Shape sh{};
  const int32_t tile = 64 * kt_for(n);
  sh.span = sum ? n : (n + tile - 1) / tile * tile;
  const int32_t need = pidx(sh.span - 1) + 1;
  // the box sum's reflected lines: extended positions [-rmax, n + rmax) at e + lead, and 32 slots
  // of slack past both ends, plain or padded (pidx); then as many rows again for H (plain), or the
  // output rows of lpad slots and a junk slot (padded)
  const int32_t rmax = std::min((max_w - 1) / 2, n - 1);
  sh.lead = 32 + rmax;
  sh.epad = !sum ? 0 : layout == 1 ? (n + 2 * rmax + 64) | 1 : (pidx(n + 2 * rmax + 64) + 1) | 1;
  if (sum) {
    sh.lpad = need | 1;
    sh.flat = 0;
  } else if (contiguous && (n & 31) == 0 && sh.span == n) {
    sh.lpad = n + n / 32;  // rows in one run: pidx(p n + i) = p lpad + pidx(i)
    sh.flat = 1;
  } else {
    sh.lpad = need | 1;
    sh.flat = 0;
  }
  if (contiguous) {
    sh.P = static_cast<int32_t>(std::min<int64_t>(a.outer, EDA_VHGW_ROWS));
  } else {
    int32_t P = static_cast<int32_t>(64 / s);  // 64-byte runs
    while (P > 1 && P / 2 >= a.inner) P /= 2;
    sh.P = P;
  }
  // max / min of a line that fits one tile keep H over the staged line: one LDS row per wave, and
  // up to 8 waves; otherwise two rows and 4 waves
  // (the box sum: the second rows per line)
  sh.wrows = sum ? 0 : sh.span == tile ? 1 : 2;
  sh.W = std::min(sum ? EDA_VHGW_SUMW : sh.wrows == 1 ? 8 : 4, sh.P);
  const auto bytes = [&](int32_t P, int32_t W) {
    if (sum && layout == 2) return (static_cast<size_t>(P) * (sh.epad + sh.lpad) + 1) * s;
    if (sum) return static_cast<size_t>(P) * 2 * sh.epad * s;
    return (static_cast<size_t>(P) * sh.lpad + static_cast<size_t>(sh.wrows * W) * (sh.lpad + 1)) * s;
  };
  const size_t budget = sum ? EDA_VHGW_SUMLDS : EDA_VHGW_LDS;
  while (bytes(sh.P, sh.W) > budget && sh.P > 1) {
    sh.P /= 2;
    sh.W = std::min(sh.W, sh.P);
  }
  sh.lds = bytes(sh.P, sh.W);
  if (sh.lds > 160 * 1024) throw std::invalid_argument("eda: the line is too long for the cube filter");
  return sh;
}

// Whether the staged kernels' LDS holds the pass (pass_shape without the exception).
bool pass_fits(const VPass& a, size_t s, bool sum, int32_t max_w, int layout = 0) {
  try {
    (void)pass_shape(a, s, sum, max_w, layout);
    return true;
  } catch (const std::invalid_argument&) {
    return false;
  }
}

template <class T, class S>
void launch_kt(int kt, dim3 grid, dim3 block, size_t lds, hipStream_t s, const VPass& a, const VPtrs<T>& p) {
  const auto go = [&](auto k) {
    if (lds > 65536)
      EDA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(lds)));
    EDA_GO(k, grid, block, lds, s, a, p);
  };
  if constexpr (std::is_same_v<S, T>) {
    go(k_vhgw<T, S, 1>);
  } else {
    switch (kt) {
      case 1: go(k_vhgw<T, S, 1>); break;
      case 2: go(k_vhgw<T, S, 2>); break;
      case 4: go(k_vhgw<T, S, 4>); break;
      case 8: go(k_vhgw<T, S, 8>); break;
      default: go(k_vhgw<T, S, 16>); break;
    }
  }
}

template <class T, class S, int KT>
void go_rows(hipStream_t s, int64_t items, const RPass& r, const VPtrs<T>& p) {
  constexpr size_t slots = 64 * KT + 2 * KT;
  const size_t per_wave = (r.il ? 4 : 2) * slots * sizeof(S);
  int W = 4;
  while (W > 1 && W * per_wave > 64 * 1024) W /= 2;
  const size_t lds = W * per_wave;
  const int64_t wgs = std::max<int64_t>(1, std::min<int64_t>((items + W - 1) / W, 8192 / W));
  auto k = k_vhgw_rows<T, S, KT>;
  if (lds > 65536)
    EDA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                                static_cast<int>(lds)));
  EDA_GO(k, dim3(static_cast<unsigned>(wgs)), dim3(64 * W), lds, s, r, p);
}

// Compute units of the current device (cached per device).
int device_cus() {
  static std::atomic<int> cache[64];
  int dev = 0;
  EDA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) dev = 0;
  int v = cache[dev].load(std::memory_order_relaxed);
  if (v <= 0) {
    EDA_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev));
    v = std::max(v, 1);
    cache[dev].store(v, std::memory_order_relaxed);
  }
  return v;
}

// The pipelined box sum takes a tile that its 16-byte reads cover in at most 8 per thread: returns
// the reads per thread it needs (4 or 8), or 0 for k_vhgw.
template <class T>
int box_reads(const VPass& a, int nthr) {
  constexpr int64_t VW = 16 / static_cast<int64_t>(sizeof(T));
  if (!a.vec || a.nf != 1) return 0;
  if (a.inner > 1 && a.inner % a.P != 0) return 0;
  const int64_t per_thread = (static_cast<int64_t>(a.P) * a.n + nthr * VW - 1) / (nthr * VW);
  if (a.inner > 1 && (a.n + nthr / (a.P / VW) - 1) / (nthr / (a.P / VW)) > per_thread) return 0;
  return per_thread <= 4 ? 4 : per_thread <= 8 ? 8 : 0;
}

template <class T, bool PAD>
void launch_box(int pf, hipStream_t s, const VPass& a, const VPtrs<T>& p, int32_t rungs, int W, size_t lds) {
  const int64_t items = a.g.batch * a.groups * rungs;
  // the candidates: 2 to 4 resident workgroups per CU (as the LDS allows); the rule's choice 4
  const int64_t cap = std::max<int64_t>(1, static_cast<int64_t>(160 * 1024 / lds));
  std::vector<int64_t> cands;
  for (int64_t pc = 2; pc <= std::min<int64_t>(4, cap); ++pc) cands.push_back(pc);
  if (cands.empty()) cands.push_back(cap);
  const int64_t rule = std::max<int64_t>(1, std::min<int64_t>(EDA_BOX_PERCU, cap));
  int fallback = static_cast<int>(cands.size()) - 1;
  for (size_t c = 0; c < cands.size(); ++c)
    if (cands[c] == rule) fallback = static_cast<int>(c);
  const auto launch = [&](int c) {
    int64_t wgs = std::min<int64_t>((items + 7) / 8 * 8, device_cus() * cands[static_cast<size_t>(c)]);
    wgs = std::max<int64_t>(8, wgs / 8 * 8);
    const auto go = [&](auto k) {
      if (lds > 65536)
        EDA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    static_cast<int>(lds)));
      VPass v = a;
      v.xcd = xcd_chunk(wgs);
      EDA_GO(k, dim3(static_cast<unsigned>(wgs)), dim3(64 * W), lds, s, v, p, rungs);
    };
    if (pf == 4) go(k_box<T, 4, PAD>);
    else go(k_box<T, 8, PAD>);
  };
  // a pass in place is not tuned (a second run would filter its own output)
  const size_t span = static_cast<size_t>(a.g.batch * a.g.count) * sizeof(T);
  const auto range = [&](const T* q, int64_t stride) {
    const char* b0 = reinterpret_cast<const char*>(q);
    return std::make_pair(b0, b0 + span + static_cast<size_t>(rungs - 1) * static_cast<size_t>(stride) * sizeof(T));
  };
  const auto ri = range(p.in[0], p.in_stride);
  const auto ro = range(p.out[0], p.out_stride);
  if (ri.first < ro.second && ro.first < ri.second) {
    launch(fallback);
    return;
  }
  int lb = 0;
  while ((int64_t{1} << lb) < items) ++lb;
  const std::string key = std::string("k_box:") + (sizeof(T) == 4 ? "f32" : "f64") + (PAD ? ":long" : ":short") +
                          ":n" + std::to_string(a.n) + ":inner" + std::to_string(a.inner) + ":P" + std::to_string(a.P) +
                          ":pf" + std::to_string(pf) + ":items2^" + std::to_string(lb);
  autotune(key, static_cast<int>(cands.size()), fallback, s, launch);
}

// The streaming pass serves max / min along a contiguous axis with n <= 2048.
bool rows_qualify(const Geom& g, int axis, bool sum) {
  const bool contiguous = axis == g.dim - 1 || g.st[axis] == 1;
  return !sum && contiguous && g.n[axis] <= 2048;
}
template <class T>
void launch_rows(hipStream_t s, const Geom& g, int axis, int nf, bool same_input, const int* ops, const VPtrs<T>& p,
                 int32_t rungs, bool il, int32_t rmin) {
  const int32_t n = g.n[axis];
  RPass r{};
  r.g = g;
  r.n = n;
  r.lines = g.count / n;
  r.items = static_cast<int64_t>(rungs) * g.batch * r.lines;
  r.nf = nf;
  r.two_in = nf == 2 && !same_input ? 1 : 0;
  r.ops[0] = ops[0];
  r.ops[1] = nf > 1 ? ops[1] : ops[0];
  r.il = il && nf == 2 && r.ops[0] == kVMax && r.ops[1] == kVMin ? 1 : 0;
  r.rmin = rmin;
  const auto al16 = [](const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; };
  r.vec = n % 4 == 0 && al16(p.in[0]) && al16(p.in[1]) && (p.in_stride * sizeof(T)) % 16 == 0 ? 1 : 0;
  const int32_t per = (n + 63) / 64;
  const int kt = per <= 1 ? 1 : per <= 2 ? 2 : per <= 4 ? 4 : per <= 8 ? 8 : per <= 16 ? 16 : 32;
  using S = okey_t<T>;
  switch (kt) {
    case 1: go_rows<T, S, 1>(s, r.items, r, p); break;
    case 2: go_rows<T, S, 2>(s, r.items, r, p); break;
    case 4: go_rows<T, S, 4>(s, r.items, r, p); break;
    case 8: go_rows<T, S, 8>(s, r.items, r, p); break;
    case 16: go_rows<T, S, 16>(s, r.items, r, p); break;
    default: go_rows<T, S, 32>(s, r.items, r, p); break;
  }
}


// ---- the streaming pass of a strided axis (windows up to 63 samples) ----------------------------------
// One wave per (rung, element, slab, group of 64 adjacent lines, piece of R outputs): lane l owns one
// line of the group and walks it sample by sample in the canonical segment order, reading one row
// of the group (64 adjacent samples, 256 bytes of float32) per step through a buffer resource whose
// row offsets live in scalar registers, so that a read or a write costs the lane no address
// arithmetic. Segments of L = 2r + 1 samples start at o + k L (the box sum's canonical offset; any
// cut gives max / min the same bits). While segment k + 1 is read, its prefix fold G meets the
// suffix fold H of segment k kept in registers: the output x = s(k+1) - r + j is
// op(H_k(s(k) + 1 + j), G(j)) for j < L - 1 and op(H_(k+1)(s(k+1)), G(L-1)) (halved for the box: the
// window is the segment) for j = L - 1; then H_(k+1) replaces H_k. G and H keep the canonical
// 32-sample chunks (L <= 32: one chunk, a plain sequential fold; L <= 63: two), so the sums are
// bitwise the staged kernels'. L is a template constant (the folds live in registers): one kernel
// takes L <= 31, another 33 <= L <= 63 (more registers), each skipping the other's elements. A wave
// reads the segments of its piece in batches of about 32 samples per lane (every read of a batch in
// flight before the first fold). No LDS and no barrier: the waves of a workgroup are independent.
// The samples the walk reads lie in [(1 - 3L) / 2, n + r + L), reflected once into the line (the
// host asks n >= 128); the pass must not run in place (the pieces of one line read each other's
// rows).
constexpr int kStreamMaxL = 31;     // the short walk
constexpr int kStreamBigL = 63;     // the long walk (two canonical chunks per box segment)

struct SPass {
  Geom g;
  int32_t n;
  int32_t lines;    // lanes per group that hold a line (64, fewer in a slab's last group)
  int64_t inner;    // adjacent lines of a slab
  int64_t outer;    // slabs per element
  int32_t cgroups;  // 64-line groups per slab
  int32_t pieces;   // pieces per line
  int32_t R;        // outputs per piece
  int32_t rmin;     // elements whose radius is below rmin are skipped (the fused cube takes them)
  int32_t mode;     // 0 max, 1 min, 2 max and min of one input (out0, out1), 3 box sum
  int32_t scale;    // multiply by fl_T(1 / prod of the clamped sides)
  int64_t items;    // waves per rung
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, uint32_t bytes) {
  // raw buffer (stride 0) of `bytes` bytes; dword 3 = the gfx9 data-format word of an untyped dword
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, static_cast<int>(bytes), 0x00020000);
}
template <class T>
__device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rs, static_cast<int>(voff), static_cast<int>(soff), 0));
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rs, static_cast<int>(voff), static_cast<int>(soff), 0));
  }
}
template <class T>
__device__ __forceinline__ void buf_st(T v, __amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(T) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int32_t, v), rs, static_cast<int>(voff), static_cast<int>(soff), 0);
  } else {
    using V2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(rs, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), rs, static_cast<int>(voff), static_cast<int>(soff), 0);
  }
}

// One segment of a walk (MODE as SPass::mode): sv the segment's L samples (order keys for max /
// min); h0 (h1) the previous segment's suffix folds of the max or the only op (of the min), replaced
// by this segment's; out: the segment has outputs, emit(j, y0, y1) for the output x0 + j, j < L
// (y1: the min of MODE 2). The box sum's folds keep the canonical chunks: a local fold restarted
// every 32 samples from the segment's start (G) or end (H), plus the carry of the chunks before
// (for L <= 32 one chunk, a plain sequential fold).
template <int MODE, int L, class S, class Emit>
__device__ __forceinline__ void seg_fold(const S* sv, S* h0, S* h1, bool out, Emit emit) {
  constexpr bool kTwo = MODE == 2;
  S g0 = sv[0], g1 = sv[0];
  S loc = sv[0], car = sv[0];  // box: the chunk's local fold and the carry of the chunks before
  const auto step_g = [&](int j) {
    if constexpr (MODE == 3 && L <= 32) {
      g0 = g0 + sv[j];
    } else if constexpr (MODE == 3) {
      if (j % 32 == 0) {
        car = j == 32 ? loc : car + loc;
        loc = sv[j];
      } else {
        loc = loc + sv[j];
      }
      g0 = j < 32 ? loc : car + loc;
    } else {
      if constexpr (MODE == 1) g0 = sv[j] < g0 ? sv[j] : g0;
      else g0 = sv[j] > g0 ? sv[j] : g0;
      if constexpr (kTwo) g1 = sv[j] < g1 ? sv[j] : g1;
    }
  };
  if (out) {
#pragma unroll
    for (int j = 0; j < L - 1; ++j) {
      if (j > 0) step_g(j);
      if constexpr (MODE == 3) {
        emit(j, h0[j + 1] + g0, g0);
      } else if constexpr (MODE == 1) {
        emit(j, h0[j + 1] < g0 ? h0[j + 1] : g0, g0);
      } else if constexpr (kTwo) {
        emit(j, h0[j + 1] > g0 ? h0[j + 1] : g0, h1[j + 1] < g1 ? h1[j + 1] : g1);
      } else {
        emit(j, h0[j + 1] > g0 ? h0[j + 1] : g0, g0);
      }
    }
    if constexpr (L > 1) step_g(L - 1);
  }
  // this segment's H (d: the distance from the segment's end)
  h0[L - 1] = sv[L - 1];
  if constexpr (kTwo) h1[L - 1] = sv[L - 1];
  S hl = sv[L - 1], hc = sv[L - 1];
#pragma unroll
  for (int j = L - 2; j >= 0; --j) {
    if constexpr (MODE == 3 && L <= 32) {
      h0[j] = h0[j + 1] + sv[j];
    } else if constexpr (MODE == 3) {
      const int d = L - 1 - j;
      if (d % 32 == 0) {
        hc = d == 32 ? hl : hc + hl;
        hl = sv[j];
      } else {
        hl = hl + sv[j];
      }
      h0[j] = d < 32 ? hl : hc + hl;
    } else if constexpr (MODE == 1) {
      h0[j] = sv[j] < h0[j + 1] ? sv[j] : h0[j + 1];
    } else {
      h0[j] = sv[j] > h0[j + 1] ? sv[j] : h0[j + 1];
    }
    if constexpr (kTwo) h1[j] = sv[j] < h1[j + 1] ? sv[j] : h1[j + 1];
  }
  // the output whose window is this segment
  if (out) {
    if constexpr (MODE == 3) emit(L - 1, (h0[0] + g0) * decltype(g0)(0.5), g0);
    else emit(L - 1, g0, g1);
  }
}

// The walk of one piece with a window of L samples (MODE as SPass::mode). src, d0, d1: the element's
// input and outputs; voff: the lane's byte offset in a row (0 for lanes without a line: they read
// the group's first line and store nothing); slab: the slab's byte offset; rowb: bytes per row
// (inner * sizeof(T)); outputs x in [x_lo, x_hi).
template <class T, int MODE, int L>
__device__ __forceinline__ void stream_walk(__amdgpu_buffer_rsrc_t src, __amdgpu_buffer_rsrc_t d0,
                                            __amdgpu_buffer_rsrc_t d1, uint32_t voff, bool live, uint32_t slab,
                                            uint32_t rowb, int32_t n, int32_t o, int32_t x_lo, int32_t x_hi, T fac,
                                            bool scale) {
  constexpr int r = (L - 1) / 2;
  constexpr int U = L >= 32 ? 1 : 32 / L;  // segments per batch of reads
  using S = std::conditional_t<MODE == 3, T, okey_t<T>>;
  // segments k1 - 1 .. k2: k1 holds x_lo + r (the first output's G), k2 holds x_hi - 1 + r
  const int32_t k1 = static_cast<int32_t>(floor_div(static_cast<int64_t>(x_lo) + r - o, L));
  const int32_t k2 = static_cast<int32_t>(floor_div(static_cast<int64_t>(x_hi) - 1 + r - o, L));
  const auto row_of = [&](int32_t e) { return e < 0 ? -1 - e : (e >= n ? 2 * n - 1 - e : e); };
  const auto load = [&](int32_t e) -> S {
    const T v = buf_ld<T>(src, voff, slab + static_cast<uint32_t>(row_of(e)) * rowb);
    if constexpr (MODE == 3) return v;
    else return okey(v);
  };
  const auto put = [&](__amdgpu_buffer_rsrc_t d, int32_t x, S y) {
    T v;
    if constexpr (MODE == 3) v = y;
    else v = okey_inv(y);
    if (scale) v = v * fac;
    if (live) buf_st<T>(v, d, voff, slab + static_cast<uint32_t>(x) * rowb);
  };
  S h0[L], h1[MODE == 2 ? L : 1];  // the suffix folds of the previous segment (max; min)
  for (int32_t k = k1 - 1; k <= k2; k += U) {
    S v[U * L];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (k + u > k2) break;
      const int32_t s = o + (k + u) * L;
#pragma unroll
      for (int j = 0; j < L; ++j) v[u * L + j] = load(s + j);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int32_t kk = k + u;
      if (kk > k2) break;
      const int32_t x0 = o + kk * L - r;  // output of step j: x0 + j
      seg_fold<MODE, L>(v + u * L, h0, h1, kk >= k1, [&](int j, S y0, S y1) {
        const int32_t x = x0 + j;
        if (x < x_lo || x >= x_hi) return;  // uniform
        put(d0, x, y0);
        if constexpr (MODE == 2) put(d1, x, y1);
      });
    }
  }
}

template <class T, int MODE, bool BIG>
__global__ void __launch_bounds__(256) k_vhgw_stream(SPass a, VPtrs<T> ptr) {
  const int64_t item = static_cast<int64_t>(blockIdx.x) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (item >= a.items) return;
  const int64_t rung = blockIdx.y;
  // item = ((b * outer + slab) * cgroups + group) * pieces + piece
  const int64_t per_b = a.outer * a.cgroups * a.pieces;
  const int64_t b = item / per_b;
  int64_t rest = item - b * per_b;
  const int64_t slab = rest / (static_cast<int64_t>(a.cgroups) * a.pieces);
  rest -= slab * a.cgroups * a.pieces;
  const int32_t grp = static_cast<int32_t>(rest / a.pieces);
  const int32_t piece = static_cast<int32_t>(rest - static_cast<int64_t>(grp) * a.pieces);
  const int32_t w = ptr.win[rung * a.g.batch + b];
  if (w < 1) return;  // uniform: the element does not run
  const int32_t n = a.n;
  const int32_t half = (w - 1) / 2;
  const int32_t r = half < n - 1 ? half : n - 1;
  if (r < a.rmin) return;  // uniform: the fused cube serves the element
  const int32_t L = 2 * r + 1;
  const int32_t o = static_cast<int32_t>((static_cast<int64_t>(n) * ((L + 1) / 2)) % L);
  const int32_t x_lo = piece * a.R;
  const int32_t x_hi = min(n, x_lo + a.R);
  const int lane = threadIdx.x & 63;
  const int64_t col = static_cast<int64_t>(grp) * 64 + lane;
  const bool live = col < a.inner;
  const uint32_t voff = live ? static_cast<uint32_t>(col * sizeof(T)) : 0u;
  const uint32_t rowb = static_cast<uint32_t>(a.inner * sizeof(T));
  const uint32_t slabb = static_cast<uint32_t>(slab * a.inner * n * sizeof(T));
  const uint32_t bytes = static_cast<uint32_t>(a.g.count * sizeof(T));
  const int64_t off = b * a.g.count;
  const __amdgpu_buffer_rsrc_t src = buf_rsrc(ptr.in[0] + rung * ptr.in_stride + off, bytes);
  const __amdgpu_buffer_rsrc_t d0 = buf_rsrc(ptr.out[0] + rung * ptr.out_stride + off, bytes);
  const __amdgpu_buffer_rsrc_t d1 = buf_rsrc(ptr.out[1] + rung * ptr.out_stride + off, bytes);
  T fac = T(1);
  if (a.scale) {
    double prod = 1.0;
    for (int ax = 0; ax < a.g.dim; ++ax) prod *= static_cast<double>(2 * min(half, a.g.n[ax] - 1) + 1);
    fac = static_cast<T>(1.0 / prod);
  }
  const bool sc = a.scale != 0;
#define EDA_STREAM_CASE(LL) \
  case LL: stream_walk<T, MODE, LL>(src, d0, d1, voff, live, slabb, rowb, n, o, x_lo, x_hi, fac, sc); break;
  // the short kernel takes windows up to 31 samples, the long one 33 to 63 (each skips the other's)
  if constexpr (!BIG) {
    switch (L) {
      EDA_STREAM_CASE(1)
      EDA_STREAM_CASE(3)
      EDA_STREAM_CASE(5)
      EDA_STREAM_CASE(7)
      EDA_STREAM_CASE(9)
      EDA_STREAM_CASE(11)
      EDA_STREAM_CASE(13)
      EDA_STREAM_CASE(15)
      EDA_STREAM_CASE(17)
      EDA_STREAM_CASE(19)
      EDA_STREAM_CASE(21)
      EDA_STREAM_CASE(23)
      EDA_STREAM_CASE(25)
      EDA_STREAM_CASE(27)
      EDA_STREAM_CASE(29)
      EDA_STREAM_CASE(31)
      default: break;
    }
  } else {
    switch (L) {
      EDA_STREAM_CASE(33)
      EDA_STREAM_CASE(35)
      EDA_STREAM_CASE(37)
      EDA_STREAM_CASE(39)
      EDA_STREAM_CASE(41)
      EDA_STREAM_CASE(43)
      EDA_STREAM_CASE(45)
      EDA_STREAM_CASE(47)
      EDA_STREAM_CASE(49)
      EDA_STREAM_CASE(51)
      EDA_STREAM_CASE(53)
      EDA_STREAM_CASE(55)
      EDA_STREAM_CASE(57)
      EDA_STREAM_CASE(59)
      EDA_STREAM_CASE(61)
      EDA_STREAM_CASE(63)
      default: break;
    }
  }
#undef EDA_STREAM_CASE
}

#ifndef EDA_CUBE2_MAXMIN
#define EDA_CUBE2_MAXMIN 0  // max + min cubes with windows up to 7 samples fused too
#endif
// ---- the fused 2-D cube (windows up to 31 samples) --------------------------------------------------
// Both passes of a 2-D max, min, max + min or box cube in one kernel, for fields of n0 >= 128 rows
// and 64 <= n1 <= 512 columns, when the cube does not run in place. A workgroup of ceil(n1 / 64)
// waves owns one (rung, element, piece of R rows). Phase 1 walks the columns in the canonical
// segment order as the streaming pass does (lane = column, rows read through a buffer resource with
// scalar row offsets) and writes each finished row of the axis-0 fold into LDS instead of HBM. After
// every batch of segments (at most 16 rows for windows up to 15 samples, L rows above) the workgroup
// synchronises and phase 2 walks the staged rows along axis 1: one lane per (row, piece of 32
// columns), rows fastest across the lanes (lanes of one step read one column of different rows:
// distinct banks, the row stride being odd), the same canonical walk over the LDS row (reflected in
// the row) writing the cube's outputs to HBM. The intermediate field never leaves the chip: the
// cube reads its input once and writes its outputs once. Max / min stay order keys between the
// passes; the box sum's folds and its final scaling are the two-pass path's, so every output is
// bitwise the two-pass result.
struct CPass {
  Geom g;
  int32_t n0, n1;  // rows (axis 0) and columns (axis 1, contiguous) of an element
  int32_t mode;    // 0 max, 1 min, 2 max and min of one input, 3 box sum
  int32_t pair;    // mode 2: max and min side by side at out[0] + 2 v (else out[0] and out[1])
  int32_t scale;   // box: multiply by fl_T(1 / window area)
  int32_t pieces;  // row pieces per element
  int32_t R;       // rows per piece
  int32_t rs;      // LDS slots per staged row (odd)
  int32_t vec;     // 16-byte output stores: n1 a multiple of their samples, outputs 16-byte aligned
};

// The canonical walk of one line in batches of U segments (segments k1 - 1 .. k2, as the streaming
// pass walks them): load(e) gives the sample at extended position e, emit(x, y0, y1) takes each
// output x in [x_lo, x_hi) in order, after(end) runs once per batch with the end of the outputs
// emitted so far. PREFETCH: the next batch's reads are issued before the batch is folded, so that
// they are in flight while after() runs. ALL: emit takes every output of the segments walked (it
// drops those outside [x_lo, x_hi) itself, without a branch).
template <int MODE, int L, int U, bool PREFETCH = false, bool ALL = false, class S, class Load, class Emit,
          class After>
__device__ __forceinline__ void walk_line(int32_t o, int32_t x_lo, int32_t x_hi, Load load, Emit emit, After after) {
  constexpr int r = (L - 1) / 2;
  const int32_t k1 = static_cast<int32_t>(floor_div(static_cast<int64_t>(x_lo) + r - o, L));
  const int32_t k2 = static_cast<int32_t>(floor_div(static_cast<int64_t>(x_hi) - 1 + r - o, L));
  S h0[L], h1[MODE == 2 ? L : 1];
  const auto fetch = [&](int32_t k, S (&v)[U * L]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (k + u > k2) break;
      const int32_t s = o + (k + u) * L;
#pragma unroll
      for (int j = 0; j < L; ++j) v[u * L + j] = load(s + j);
    }
  };
  S v[U * L];
  if constexpr (PREFETCH) fetch(k1 - 1, v);
  for (int32_t k = k1 - 1; k <= k2; k += U) {
    // PREFETCH: the next batch's reads go out before this batch is folded and after() runs
    S nv[PREFETCH ? U * L : 1];
    if constexpr (PREFETCH) {
      if (k + U <= k2) fetch(k + U, nv);
    } else {
      fetch(k, v);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int32_t kk = k + u;
      if (kk > k2) break;
      const int32_t x0 = o + kk * L - r;
      seg_fold<MODE, L>(v + u * L, h0, h1, kk >= k1, [&](int j, S y0, S y1) {
        const int32_t x = x0 + j;
        if constexpr (ALL) {
          emit(x, y0, y1);  // every output; emit keeps those in [x_lo, x_hi)
        } else {
          if (x < x_lo || x >= x_hi) return;
          emit(x, y0, y1);
        }
      });
    }
    const int32_t kl = k + U - 1 < k2 ? k + U - 1 : k2;
    const int32_t end = o + (kl + 1) * L - r;
    after(end < x_hi ? end : x_hi);
    if constexpr (PREFETCH) {
#pragma unroll
      for (int j = 0; j < U * L; ++j) v[j] = nv[j];
    }
  }
}

// Rows per batch of the fused cube: about 16 (8 for max + min, which stages three regions and
// takes windows up to 7 samples), or L.
template <int MODE, int L>
constexpr int cube_u() {
  return L >= (MODE == 2 ? 8 : 16) ? 1 : (MODE == 2 ? 8 : 16) / L;
}
template <int MODE, int L>
constexpr int cube_cap() {
  return cube_u<MODE, L>() * L;
}

// Reflected samples staged on each side of a row: the phase-2 walk reads [-(3L - 1) / 2, n1 + (3L - 1) / 2).
__host__ __device__ constexpr int cube2_halo(int L) { return L <= 15 ? 24 : 48; }

// VS consecutive values of T (VS * sizeof(T) = 4, 8 or 16 bytes) in one buffer store.
template <class T, int VS>
__device__ __forceinline__ void buf_st_vec(const T (&v)[VS], __amdgpu_buffer_rsrc_t rs, uint32_t voff) {
  constexpr int kB = VS * static_cast<int>(sizeof(T));
  if constexpr (kB == 16) {
    using V4 = decltype(__builtin_amdgcn_raw_buffer_load_b128(rs, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V4, v), rs, static_cast<int>(voff), 0, 0);
  } else if constexpr (kB == 8) {
    using V2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(rs, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), rs, static_cast<int>(voff), 0, 0);
  } else {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int32_t, v), rs, static_cast<int>(voff), 0, 0);
  }
}

// A workgroup barrier that orders LDS only: global reads in flight stay in flight across it (a
// plain __syncthreads waits for them).
__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <class T, int MODE, int L>
__device__ __forceinline__ void cube2_run(const CPass& a, const VPtrs<T>& ptr, int64_t b, int64_t rung, int32_t piece,
                                          unsigned char* smem) {
  using S = std::conditional_t<MODE == 3, T, okey_t<T>>;
  constexpr int U = cube_u<MODE, L>();
  constexpr int CAP = U * L;            // rows staged per batch at most
  constexpr int PW = CAP <= 16 ? 16 : 32;  // columns per phase-2 piece
  constexpr int HB = cube2_halo(L);        // reflected samples staged on each side of a row
  const int32_t n0 = a.n0, n1 = a.n1, rs = a.rs;
  // LDS: the staged rows B0 (B1: the min of max + min), sample e of a row at HB + e for e in
  // [-HB, n1 + HB) (the reflection of the row: phase 2 reads no index twice), and the output rows O
  // (output x at x, slot n1 a junk slot for the outputs a piece drops)
  S* B0 = reinterpret_cast<S*>(smem);
  S* B1 = B0 + CAP * rs;
  S* O = MODE == 2 ? B1 + CAP * rs : B1;
  const int tid = threadIdx.x;
  const int nthr = blockDim.x;
  const bool live = tid < n1;
  const uint32_t voff = live ? static_cast<uint32_t>(tid) * sizeof(T) : 0u;
  const uint32_t rowb = static_cast<uint32_t>(n1) * sizeof(T);
  const int64_t off = b * a.g.count;
  const int pf = a.pair ? 2 : 1;
  const __amdgpu_buffer_rsrc_t src =
      buf_rsrc(ptr.in[0] + rung * ptr.in_stride + off, static_cast<uint32_t>(a.g.count * sizeof(T)));
  const uint32_t obytes = static_cast<uint32_t>(a.g.count * pf * sizeof(T));
  const __amdgpu_buffer_rsrc_t d0 = buf_rsrc(ptr.out[0] + rung * ptr.out_stride + off * pf, obytes);
  const __amdgpu_buffer_rsrc_t d1 = buf_rsrc(ptr.out[1] + rung * ptr.out_stride + off * pf, obytes);
  T fac = T(1);
  if (a.scale) fac = static_cast<T>(1.0 / (static_cast<double>(L) * static_cast<double>(L)));
  const int32_t o0 = static_cast<int32_t>((static_cast<int64_t>(n0) * ((L + 1) / 2)) % L);
  const int32_t o1 = static_cast<int32_t>((static_cast<int64_t>(n1) * ((L + 1) / 2)) % L);
  const int32_t x_lo = piece * a.R;
  const int32_t x_hi = min(n0, x_lo + a.R);
  // phase 2's lane: row i2 of the batch (fastest), piece p2 of PW columns
  const int32_t i2 = tid % CAP;
  const int32_t c_lo = (tid / CAP) * PW;
  const int32_t c_hi = min(n1, c_lo + PW);
  // the waves holding the columns whose reflections the halo stages
  const bool edge_wave = (tid & ~63) < HB || (tid | 63) >= n1 - HB;
  // the write-out: VS consecutive samples per thread and store (16 bytes where the rows allow)
  const auto write_out = [&](int32_t xr0, int32_t nr, auto vs_tag) {
    constexpr int VS = decltype(vs_tag)::value;
    const int32_t per = n1 / VS;
    const float inv = 1.0f / static_cast<float>(per);
    for (int32_t q = tid; q < nr * per; q += nthr) {
      const int32_t i = div_small(q, per, inv);
      const int32_t c = (q - i * per) * VS;
      const S* ro = O + i * rs + c;
      const uint32_t at = (static_cast<uint32_t>(xr0 + i) * static_cast<uint32_t>(n1) + static_cast<uint32_t>(c)) *
                          static_cast<uint32_t>(pf) * sizeof(T);
      if constexpr (MODE == 2) {
        const S* rm = B0 + i * rs + c;  // the min rows
        if (a.pair) {
          if constexpr (2 * VS * sizeof(T) <= 16) {
            T v[2 * VS];
#pragma unroll
            for (int k = 0; k < VS; ++k) {
              v[2 * k] = okey_inv(ro[k]);
              v[2 * k + 1] = okey_inv(rm[k]);
            }
            buf_st_vec<T, 2 * VS>(v, d0, at);
          }
        } else {
          T v0[VS], v1[VS];
#pragma unroll
          for (int k = 0; k < VS; ++k) {
            v0[k] = okey_inv(ro[k]);
            v1[k] = okey_inv(rm[k]);
          }
          buf_st_vec<T, VS>(v0, d0, at);
          buf_st_vec<T, VS>(v1, d1, at);
        }
      } else {
        T v[VS];
#pragma unroll
        for (int k = 0; k < VS; ++k) {
          if constexpr (MODE == 3) v[k] = a.scale ? ro[k] * fac : ro[k];
          else v[k] = okey_inv(ro[k]);
        }
        buf_st_vec<T, VS>(v, d0, at);
      }
    }
  };
  int32_t flushed = x_lo;  // rows before it are written
  const auto rows = [&](int32_t end) {
    if (end <= flushed) return;  // uniform
    lds_sync();
    const int32_t nr = end - flushed;
    const bool act = i2 < nr && c_lo < n1;
    const auto none = [](int32_t) {};
    if (act) {
      S* orow = O + i2 * rs;
      const S* row0 = B0 + i2 * rs + HB;
      walk_line<MODE == 2 ? 0 : MODE, L, U, false, true, S>(
          o1, c_lo, c_hi, [&](int32_t e) { return row0[e]; },
          [&](int32_t x, S y0, S) { orow[x >= c_lo && x < c_hi ? x : n1] = y0; }, none);
    }
    if constexpr (MODE == 2) {
      lds_sync();  // B0 is read; the min rows go over it
      if (act) {
        S* orow = B0 + i2 * rs;
        const S* row1 = B1 + i2 * rs + HB;
        walk_line<1, L, U, false, true, S>(
            o1, c_lo, c_hi, [&](int32_t e) { return row1[e]; },
            [&](int32_t x, S y0, S) { orow[x >= c_lo && x < c_hi ? x : n1] = y0; }, none);
      }
    }
    lds_sync();
    // 16-byte stores where the rows and the outputs allow them (a.vec, from the host)
    constexpr int kVec = static_cast<int>(16 / sizeof(T));
    if (a.vec) {
      if (a.pair) write_out(flushed, nr, std::integral_constant<int, kVec / 2>{});
      else write_out(flushed, nr, std::integral_constant<int, kVec>{});
    } else {
      write_out(flushed, nr, std::integral_constant<int, 1>{});
    }
    // max + min wrote its min rows over B0, which the next batch stages into; otherwise the next
    // batch writes B0 (read before the last barrier) and O only after its first barrier
    if constexpr (MODE == 2) lds_sync();
    flushed = end;
  };
  const auto row_of0 = [&](int32_t e) { return e < 0 ? -1 - e : (e >= n0 ? 2 * n0 - 1 - e : e); };
  walk_line<MODE, L, U, true, false, S>(
      o0, x_lo, x_hi,
      [&](int32_t e) -> S {
        const T v = buf_ld<T>(src, voff, static_cast<uint32_t>(row_of0(e)) * rowb);
        if constexpr (MODE == 3) return v;
        else return okey(v);
      },
      [&](int32_t x, S y0, S y1) {
        if (!live) return;
        S* r0 = B0 + (x - flushed) * rs + HB;
        r0[tid] = y0;
        if constexpr (MODE == 2) B1[(x - flushed) * rs + HB + tid] = y1;
        if (edge_wave) {  // the row's reflections: e = -1 - c for c < HB, e = 2 n1 - 1 - c for c >= n1 - HB
          if (tid < HB) {
            r0[-1 - tid] = y0;
            if constexpr (MODE == 2) B1[(x - flushed) * rs + HB - 1 - tid] = y1;
          }
          if (tid >= n1 - HB) {
            r0[2 * n1 - 1 - tid] = y0;
            if constexpr (MODE == 2) B1[(x - flushed) * rs + HB + 2 * n1 - 1 - tid] = y1;
          }
        }
      },
      rows);
}

template <class T, int MODE, bool MID>
__global__ void __launch_bounds__(512) k_cube2(CPass a, VPtrs<T> ptr) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int64_t b = blockIdx.x / a.pieces;
  const int32_t piece = static_cast<int32_t>(blockIdx.x - b * a.pieces);
  const int64_t rung = blockIdx.y;
  const int32_t w = ptr.win[rung * a.g.batch + b];
  if (w < 1) return;  // uniform: the element does not run
  // n0 >= 128 and n1 >= 64: the radius is (w - 1) / 2 on both axes for every window this kernel takes
  const int32_t L = 2 * ((w - 1) / 2) + 1;
#define EDA_CUBE2_CASE(LL) \
  case LL: cube2_run<T, MODE, LL>(a, ptr, b, rung, piece, smem); break;
  if constexpr (MODE == 2) {
    if constexpr (!MID) {
      switch (L) {  // max + min: windows up to 7 samples (the two-pass path takes the longer ones)
        EDA_CUBE2_CASE(1)
        EDA_CUBE2_CASE(3)
        EDA_CUBE2_CASE(5)
        EDA_CUBE2_CASE(7)
        default: break;
      }
    }
  } else if constexpr (!MID) {
    switch (L) {
      EDA_CUBE2_CASE(1)
      EDA_CUBE2_CASE(3)
      EDA_CUBE2_CASE(5)
      EDA_CUBE2_CASE(7)
      EDA_CUBE2_CASE(9)
      EDA_CUBE2_CASE(11)
      EDA_CUBE2_CASE(13)
      EDA_CUBE2_CASE(15)
      default: break;  // another launch's element
    }
  } else {
    switch (L) {
      EDA_CUBE2_CASE(17)
      EDA_CUBE2_CASE(19)
      EDA_CUBE2_CASE(21)
      EDA_CUBE2_CASE(23)
      EDA_CUBE2_CASE(25)
      EDA_CUBE2_CASE(27)
      EDA_CUBE2_CASE(29)
      EDA_CUBE2_CASE(31)
      default: break;
    }
  }
#undef EDA_CUBE2_CASE
}

// The streaming pass takes the elements of a strided axis whose windows are at most 63 samples when
// the line has at least 128 samples, the pass does not run in place, one element's field is under
// 4 GiB (32-bit buffer offsets), and its operators are one of max, min, box or max + min of one
// input.
template <class T>
bool stream_qualifies(const Geom& g, int axis, int nf, bool same_input, const int* ops, const VPtrs<T>& p) {
  if (axis == g.dim - 1 || g.st[axis] == 1) return false;
  const int32_t n = g.n[axis];
  if (n < 128) return false;
  if (static_cast<double>(g.count) * sizeof(T) >= 4294967295.0) return false;
  if (nf == 2 && !(same_input && ops[0] == kVMax && ops[1] == kVMin)) return false;
  if (nf == 1 && ops[0] != kVMax && ops[0] != kVMin && ops[0] != kVSum) return false;
  // in place: an output field is an input field (their rung strides overlap)
  for (int f = 0; f < nf; ++f)
    for (int i = 0; i < nf; ++i)
      if (p.out[f] == p.in[i]) return false;
  return true;
}

template <class T>
void launch_stream(hipStream_t s, const Geom& g, int axis, int nf, const int* ops, bool scale, const VPtrs<T>& p,
                   int32_t rungs, int32_t max_w, bool big, int32_t rmin) {
  SPass a{};
  a.g = g;
  a.rmin = rmin;
  a.n = g.n[axis];
  a.inner = g.st[axis];
  a.outer = g.count / (static_cast<int64_t>(a.n) * a.inner);
  a.cgroups = static_cast<int32_t>((a.inner + 63) / 64);
  a.scale = scale ? 1 : 0;
  a.mode = nf == 2 ? 2 : (ops[0] == kVSum ? 3 : (ops[0] == kVMin ? 1 : 0));
  // pieces: enough waves to fill the chip (about 8 per SIMD over the launch), each piece at least
  // 4 windows long so that the walk's lead-in (one segment and r samples) stays a small part of it
  const int64_t lines = static_cast<int64_t>(rungs) * g.batch * a.outer * a.cgroups;
  const int64_t want = static_cast<int64_t>(device_cus()) * 4 * 8;
  const int32_t Lm = 2 * std::min((max_w - 1) / 2, a.n - 1) + 1;
  // the candidate piece counts (powers of two, each piece at least 32 samples and 2 windows long),
  // and the rule's choice: enough waves for about 8 per SIMD, pieces of at least 64 samples and 4
  // windows
  std::vector<int32_t> cands;
  for (int32_t pc = 1; pc <= 32 && a.n / pc >= std::max(32, 2 * Lm); pc *= 2) cands.push_back(pc);
  int32_t rule = 1;
  while (lines * rule < want && a.n / (2 * rule) >= std::max(64, 4 * Lm)) rule *= 2;
  int fallback = -1;
  for (size_t c = 0; c < cands.size(); ++c)
    if (cands[c] == rule) fallback = static_cast<int>(c);
  if (fallback < 0) {  // long lines of few columns: the rule splits them further than the list goes
    cands.push_back(rule);
    fallback = static_cast<int>(cands.size()) - 1;
  }
  const auto launch = [&](int c) {
    SPass v = a;
    v.pieces = cands[static_cast<size_t>(c)];
    v.R = (v.n + v.pieces - 1) / v.pieces;
    v.items = static_cast<int64_t>(g.batch) * v.outer * v.cgroups * v.pieces;
    if ((v.items + 3) / 4 > 0x7fffffff) throw std::invalid_argument("eda: too many waves for one pass");
    const dim3 grid(static_cast<unsigned>((v.items + 3) / 4), static_cast<unsigned>(rungs));
    const auto go = [&](auto big_tag) {
      constexpr bool kBig = decltype(big_tag)::value;
      switch (v.mode) {
        case 0: EDA_GO((k_vhgw_stream<T, 0, kBig>), grid, dim3(256), 0, s, v, p); break;
        case 1: EDA_GO((k_vhgw_stream<T, 1, kBig>), grid, dim3(256), 0, s, v, p); break;
        case 2: EDA_GO((k_vhgw_stream<T, 2, kBig>), grid, dim3(256), 0, s, v, p); break;
        default: EDA_GO((k_vhgw_stream<T, 3, kBig>), grid, dim3(256), 0, s, v, p); break;
      }
    };
    if (big) go(std::true_type{});
    else go(std::false_type{});
  };
  // the key: the kernel, the line and the parallelism (the lines of the launch, to a power of two)
  int lb = 0;
  while ((int64_t{1} << lb) < lines) ++lb;
  const std::string key = std::string("k_vhgw_stream:") + (sizeof(T) == 4 ? "f32" : "f64") + ":mode" +
                          std::to_string(a.mode) + (big ? ":big" : ":short") + ":n" + std::to_string(a.n) +
                          ":inner" + std::to_string(a.inner) + ":lines2^" + std::to_string(lb);
  autotune(key, static_cast<int>(cands.size()), fallback, s, launch);
}

// ---- the fused 2-D cube's launches ----
// LDS slots per staged row: the row, its two halos and a junk slot, odd.
inline int32_t cube2_rs(int32_t n1, bool mid) { return (n1 + 2 * cube2_halo(mid ? 31 : 15) + 1) | 1; }

// LDS of one workgroup: the staged rows of a batch (16, or 31 for windows 17 .. 31), per output field.
template <class T>
size_t cube2_lds(int32_t n1, int mode, bool mid) {
  // the staged rows (two regions for max + min) and the output rows, CAP rows each
  const int regions = mode == 2 ? 3 : 2;
  const int cap = mid ? 31 : (mode == 2 ? 8 : 16);
  return static_cast<size_t>(cap) * static_cast<size_t>(cube2_rs(n1, mid)) * static_cast<size_t>(regions) * sizeof(T);
}

// The fused cube takes a 2-D launch of n0 >= 128 rows and 64 <= n1 <= 512 columns whose input lies
// outside every output field, when 32-bit offsets reach an element's outputs and the LDS holds a
// batch of the longer class.
template <class T>
bool cube2_qualifies(const Geom& g, int mode, bool pair, const VPtrs<T>& p, int32_t rungs) {
  if (g.dim != 2 || g.n[0] < 128 || g.n[1] < 64 || g.n[1] > 512 || g.st[1] != 1) return false;
  const int pf = pair ? 2 : 1;
  if (static_cast<double>(g.count) * pf * sizeof(T) >= 4294967295.0) return false;
  if (cube2_lds<T>(g.n[1], mode, false) > 160 * 1024) return false;
  const auto range = [&](const T* q, int64_t per, int64_t stride) {
    const char* b0 = reinterpret_cast<const char*>(q);
    return std::make_pair(b0, b0 + (static_cast<size_t>(per) + static_cast<size_t>(rungs - 1) * stride) * sizeof(T));
  };
  const auto ri = range(p.in[0], g.batch * g.count, p.in_stride);
  for (int f = 0; f < (mode == 2 && !pair ? 2 : 1); ++f) {
    const auto ro = range(p.out[f], g.batch * g.count * pf, p.out_stride);
    if (ri.first < ro.second && ro.first < ri.second) return false;
  }
  return true;
}

template <class T>
void launch_cube2(hipStream_t s, const Geom& g, int mode, bool pair, bool scale, const VPtrs<T>& p, int32_t rungs,
                  bool mid) {
  CPass a{};
  a.g = g;
  a.n0 = g.n[0];
  a.n1 = g.n[1];
  a.mode = mode;
  a.pair = pair ? 1 : 0;
  a.scale = scale ? 1 : 0;
  a.rs = cube2_rs(a.n1, mid);
  const size_t lds = cube2_lds<T>(a.n1, mode, mid);
  {
    const int64_t vs = 16 / static_cast<int64_t>(sizeof(T)) / (pair ? 2 : 1);
    const auto al16 = [](const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; };
    a.vec = a.n1 % vs == 0 && al16(p.out[0]) && al16(p.out[1]) && (p.out_stride * sizeof(T)) % 16 == 0 &&
                    (g.count * (pair ? 2 : 1) * sizeof(T)) % 16 == 0
                ? 1
                : 0;
  }
  const int waves = (a.n1 + 63) / 64;
  // the candidate row pieces per element (powers of two, at least 32 rows each) and the rule's
  // choice: the fewest that give two workgroups per CU over the launch
  std::vector<int32_t> cands;
  for (int32_t pc = 1; pc <= 16 && a.n0 / pc >= 32; pc *= 2) cands.push_back(pc);
  int32_t rule = 1;
  const int64_t want = 2 * static_cast<int64_t>(device_cus());
  while (g.batch * rungs * rule < want && a.n0 / (2 * rule) >= 32 && rule < 16) rule *= 2;
  int fallback = 0;
  for (size_t c = 0; c < cands.size(); ++c)
    if (cands[c] == rule) fallback = static_cast<int>(c);
  const auto launch = [&](int c) {
    CPass v = a;
    v.pieces = cands[static_cast<size_t>(c)];
    v.R = (v.n0 + v.pieces - 1) / v.pieces;
    const int64_t wgs = g.batch * v.pieces;
    if (wgs > 0x7fffffff) throw std::invalid_argument("eda: too many workgroups for one cube");
    const dim3 grid(static_cast<unsigned>(wgs), static_cast<unsigned>(rungs));
    const auto go = [&](auto k) {
      if (lds > 65536)
        EDA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    static_cast<int>(lds)));
      EDA_GO(k, grid, dim3(64 * waves), lds, s, v, p);
    };
    const auto by_mode = [&](auto mid_tag) {
      constexpr bool kMid = decltype(mid_tag)::value;
      switch (mode) {
        case 0: go(k_cube2<T, 0, kMid>); break;
        case 1: go(k_cube2<T, 1, kMid>); break;
        case 2: go(k_cube2<T, 2, kMid>); break;
        default: go(k_cube2<T, 3, kMid>); break;
      }
    };
    if (mid) by_mode(std::true_type{});
    else by_mode(std::false_type{});
  };
  int lb = 0;
  while ((int64_t{1} << lb) < g.batch * rungs) ++lb;
  const std::string key = std::string("k_cube2:") + (sizeof(T) == 4 ? "f32" : "f64") + ":mode" + std::to_string(mode) +
                          (pair ? ":pair" : "") + (mid ? ":mid" : ":small") + ":n" + std::to_string(a.n0) + "x" +
                          std::to_string(a.n1) + ":elems2^" + std::to_string(lb);
  autotune(key, static_cast<int>(cands.size()), fallback, s, launch);
}

}  // namespace

// ---- lines too long for LDS (the whole pass in global memory) ---------------------------------------
// The canonical vHGW order of a line of n samples and radius r = min((w - 1) / 2, n - 1): segments
// of wa = 2r + 1 positions of the reflected extension starting at o + k wa, o = (n (wa + 1) / 2)
// mod wa; inside a segment the prefix fold G runs in 32-sample chunks (a sequential fold in the
// chunk, the chunk totals folded into a carry in order), the suffix fold H is G of the reversed
// line, y(i) = op(H(i - r), G(i + r)) (the sum of an aligned window 0.5 (H + G)). Here each of its
// three steps is a kernel over global memory: the chunk folds (one thread per chunk: the running
// fold from the chunk start, kept at the positions the outputs read, and the chunk total), the
// carries (one thread per segment, in order), and every output from two carries and two running
// folds. Used for passes whose line does not fit the LDS of the staged kernels.
struct LongPass {
  Geom g;
  int32_t n, nf, ops[2], scale, rungs;
  int64_t step, inner, lines, cap;  // cap: chunk slots per line and direction (n + 2 rmax)
};

template <class T>
__device__ __forceinline__ T long_apply(int op, T a, T b) {
  return op == kVMax ? max_c(a, b) : (op == kVMin ? min_c(a, b) : a + b);
}

struct LongGeo {
  int32_t r, wa, o, k_first, nseg, cps;
};
__device__ __forceinline__ LongGeo long_geo(int32_t n, int32_t w) {
  LongGeo q;
  q.r = min((w - 1) / 2, n - 1);
  q.wa = 2 * q.r + 1;
  q.o = static_cast<int32_t>((static_cast<int64_t>(n) * ((q.wa + 1) / 2)) % q.wa);
  q.k_first = static_cast<int32_t>(floor_div(static_cast<int64_t>(q.r) - q.o, q.wa));
  const int32_t k_last = static_cast<int32_t>(floor_div(static_cast<int64_t>(n) - 1 + q.r - q.o, q.wa));
  q.nseg = k_last - q.k_first + 1;
  q.cps = (q.wa + 31) / 32;
  return q;
}
// the value at extended position e of the line (rev: of the reversed line)
template <class T>
__device__ __forceinline__ T long_at(const T* line, int64_t step, int32_t n, int32_t e, bool rev) {
  int32_t i = reflect_index(e, n);
  if (rev) i = n - 1 - i;
  return line[static_cast<int64_t>(i) * step];
}
// the base of line `li` of element b along the pass's axis
__device__ __forceinline__ int64_t long_line_base(const LongPass& a, int64_t b, int64_t li) {
  if (a.inner == 1) return b * a.g.count + li * a.n;
  const int64_t o = li / a.inner, c = li - o * a.inner;
  return b * a.g.count + o * static_cast<int64_t>(a.n) * a.inner + c;
}

// The long pass's buffers: the inputs, the chunk totals and carries (slot), the running folds (pre).
template <class T>
struct LongArgs {
  LongPass a;
  const int32_t* win;
  const T* in[2];
  T* out[2];
  int64_t in_stride, out_stride, B;
  T* ws;
  T* pre;
  // field f, direction dir (0 G, 1 the reversed line), total (0) / carry (1), rung j, element b, line li
  __device__ __forceinline__ int64_t slot_of(int f, int dir, int which, int64_t j, int64_t b, int64_t li) const {
    return ((((static_cast<int64_t>(f) * 2 + dir) * 2 + which) * a.rungs + j) * B + b) * a.lines * a.cap + li * a.cap;
  }
  __device__ __forceinline__ int64_t pre_of(int f, int dir, int64_t j, int64_t b, int64_t li) const {
    return (((static_cast<int64_t>(f) * 2 + dir) * a.rungs + j) * B + b) * a.g.count + li * a.n;
  }
};

// A grid row's task, decoded once per workgroup into LDS: rung j, element b, line li (and the
// direction for the fold and carry steps), the window's geometry and the line's base offset.
struct LongTask {
  int64_t j, b, li, base;
  int32_t dir, w;
  LongGeo G;
};
template <class T>
__device__ __forceinline__ void long_task(const LongArgs<T>& A, int64_t t, bool with_dir, LongTask* out) {
  LongTask k;
  int64_t q = t;
  k.dir = 0;
  if (with_dir) {
    k.dir = static_cast<int32_t>(q & 1);
    q >>= 1;
  }
  k.li = q % A.a.lines;
  q /= A.a.lines;
  k.b = q % A.B;
  k.j = q / A.B;
  k.w = A.win[k.j * A.B + k.b];
  if (k.w >= 1) k.G = long_geo(A.a.n, k.w);
  k.base = long_line_base(A.a, k.b, k.li);
  *out = k;
}

// Step 1: one wave per 64 consecutive chunks of a task (the waves of a grid row stride over them):
// the chunks tile the extended line, so their samples are one range of at most 2048 positions, which
// the wave loads coalesced into LDS (a pad word every 32 keeps the lanes' chunk walks on distinct
// banks); each lane then folds its chunk in order there, keeping the running fold in place, and the
// wave stores the running folds at [r, n - 1 + r] coalesced, and the chunk totals.
constexpr int kFoldRange = 64 * 32;
__device__ __forceinline__ int fold_pad(int p) { return p + (p >> 5); }

template <class T>
constexpr int fold_waves() {
  return sizeof(T) == 8 ? 2 : 4;  // 34 KiB of LDS per workgroup either way
}
// The pass's operators as one compile-time mode: max, min, box sum (one field each) or max and min
// (two fields); field f's operator.
constexpr int kLongMax = 0, kLongMin = 1, kLongSum = 2, kLongMaxMin = 3;
template <int MODE>
__host__ __device__ constexpr int long_nf() {
  return MODE == kLongMaxMin ? 2 : 1;
}
template <int MODE>
__host__ __device__ constexpr int long_op(int f) {
  return MODE == kLongMaxMin ? (f == 0 ? kVMax : kVMin) : (MODE == kLongMax ? kVMax : (MODE == kLongMin ? kVMin : kVSum));
}
// the value at extended position e of a line with unit stride (the contiguous axis) or any stride
template <class T>
__device__ __forceinline__ T long_at_s(const T* line, int64_t step, int32_t n, int32_t e, bool rev) {
  int32_t i = reflect_index(e, n);
  if (rev) i = n - 1 - i;
  return step == 1 ? line[i] : line[static_cast<int64_t>(i) * step];
}

template <class T, int MODE>
__global__ void __launch_bounds__(256) k_long_fold(LongArgs<T> A) {
  constexpr int kW = fold_waves<T>();
  __shared__ LongTask st;
  __shared__ T buf[kW][kFoldRange + kFoldRange / 32];
  const LongPass& a = A.a;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  T* lds = buf[wave];
  const int64_t tasks = static_cast<int64_t>(a.rungs) * A.B * a.lines * 2;
  for (int64_t t = blockIdx.y; t < tasks; t += gridDim.y) {
    if (threadIdx.x == 0) long_task(A, t, true, &st);
    __syncthreads();
    const LongTask k = st;
    __syncthreads();
    if (k.w < 1) continue;
    const LongGeo& G = k.G;
    const int32_t nchunks = G.nseg * G.cps;
    const int32_t last = a.n - 1 + G.r;  // the folds stop at n - 1 + r
    const bool rev = k.dir == 1;
    // chunk c's first and last position (e0 > last: a chunk past the folds' end)
    const auto span = [&](int32_t c, int32_t* e0, int32_t* e1) {
      const int32_t seg = c / G.cps, qc = c - seg * G.cps;
      const int32_t s0 = G.o + (G.k_first + seg) * G.wa;
      *e0 = s0 + 32 * qc;
      *e1 = min(min(*e0 + 31, s0 + G.wa - 1), last);
    };
    for (int32_t c0 = (blockIdx.x * kW + wave) * 64; c0 < nchunks; c0 += gridDim.x * kW * 64) {
      const int32_t c = c0 + lane;
      int32_t e0 = 0, e1 = -1;
      if (c < nchunks) span(c, &e0, &e1);
      const bool live = c < nchunks && e0 <= last;
      int32_t lo, hi;  // the wave's range: its first chunk's start to its last live chunk's end
      span(c0, &lo, &hi);
      hi = live ? e1 : lo - 1;
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1) hi = max(hi, __shfl_xor(hi, d, 64));
      const int cnt = hi - lo + 1;  // <= 2048
      if (cnt <= 0) continue;       // uniform over the wave
      const int rounds = (cnt + 63) / 64;
      const int b0 = e0 - lo, len = e1 - e0 + 1;
#pragma unroll
      for (int f = 0; f < long_nf<MODE>(); ++f) {
        const int op = long_op<MODE>(f);
        const T* line = A.in[f] + k.j * A.in_stride + k.base;
        T* P = A.pre + A.pre_of(f, k.dir, k.j, k.b, k.li) - G.r;  // P[e] for e in [r, n - 1 + r]
        // coalesced loads, 8 in flight at a time (clamped reads: always valid)
        for (int u0 = 0; u0 < rounds; u0 += 8) {
          T v[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) v[u] = long_at_s(line, a.step, a.n, lo + min(lane + 64 * (u0 + u), cnt - 1), rev);
#pragma unroll
          for (int u = 0; u < 8; ++u)
            if (lane + 64 * (u0 + u) < cnt) lds[fold_pad(lane + 64 * (u0 + u))] = v[u];
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        if (live) {
          T acc = lds[fold_pad(b0)];
          for (int u0 = 1; u0 < len; u0 += 8) {
            T v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = lds[fold_pad(b0 + min(u0 + u, len - 1))];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              if (u0 + u < len) {
                acc = long_apply(op, acc, v[u]);
                lds[fold_pad(b0 + u0 + u)] = acc;
              }
            }
          }
          A.ws[A.slot_of(f, k.dir, 0, k.j, k.b, k.li) + c] = acc;
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        for (int u = 0; u < rounds; ++u) {
          const int q = lane + 64 * u;
          if (q < cnt && lo + q >= G.r) P[lo + q] = lds[fold_pad(q)];
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
      }
    }
  }
}

// The fold through the chunk totals tot[1 .. chunks - 1] of one segment in order, from carry =
// tot[0]: car[q] = the carry before chunk q. One thread runs the chain (it is sequential in the
// canonical order): 64 totals at a time into registers (the loads in flight together), the carries
// over them in place, 64 stores.
template <class T, int OP>
__device__ __forceinline__ void carry_run(const T* tot, T* car, int32_t chunks) {
  T carry = tot[0];
  int32_t q0 = 1;
  for (; q0 + 64 <= chunks; q0 += 64) {
    T t[64];
#pragma unroll
    for (int u = 0; u < 64; ++u) t[u] = tot[q0 + u];
#pragma unroll
    for (int u = 0; u < 64; ++u) {
      const T x = t[u];
      t[u] = carry;
      carry = long_apply(OP, carry, x);
    }
#pragma unroll
    for (int u = 0; u < 64; ++u) car[q0 + u] = t[u];
  }
  for (; q0 < chunks; ++q0) {
    car[q0] = carry;
    carry = long_apply(OP, carry, tot[q0]);
  }
}

// The same carries for max / min, which do not depend on the order of the fold: one wave per
// segment, 64 totals at a time through a wave prefix scan.
template <class T, int OP>
__device__ __forceinline__ void carry_scan(const T* tot, T* car, int32_t chunks, int lane) {
  T carry = tot[0];
  for (int32_t q0 = 1; q0 < chunks; q0 += 64) {
    const int32_t qi = q0 + lane;
    T x = tot[qi < chunks ? qi : q0];  // lanes past the end repeat a total: only the ones before read it
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const T u = __shfl_up(x, d, 64);
      if (lane >= d) x = long_apply(OP, u, x);
    }
    const T before = __shfl_up(x, 1, 64);
    if (qi < chunks) car[qi] = lane == 0 ? carry : long_apply(OP, carry, before);
    const int m = min(64, chunks - q0);
    carry = long_apply(OP, carry, __shfl(x, m - 1, 64));
  }
}

// Step 2: the carry before each chunk of a segment (chunk 0 has none), for the segments of two
// chunks or more (wa > 32): one thread per segment through carry_run, or for max / min over
// segments of more than 64 chunks one wave per segment through carry_scan (the workgroups of a grid
// row stride over the segments of one task).
template <class T, int MODE>
__global__ void __launch_bounds__(256) k_long_carry(LongArgs<T> A) {
  __shared__ LongTask st;
  const LongPass& a = A.a;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t tasks = static_cast<int64_t>(a.rungs) * A.B * a.lines * 2;
  for (int64_t t = blockIdx.y; t < tasks; t += gridDim.y) {
    if (threadIdx.x == 0) long_task(A, t, true, &st);
    __syncthreads();
    const LongTask k = st;
    __syncthreads();
    if (k.w < 1 || k.G.cps < 2) continue;
    const LongGeo& G = k.G;
    const int32_t last = a.n - 1 + G.r;
    if (MODE != kLongSum && G.cps > 64) {
      for (int32_t seg = blockIdx.x * 4 + wave; seg < G.nseg; seg += gridDim.x * 4) {
        const int32_t s0 = G.o + (G.k_first + seg) * G.wa;
        const int32_t chunks = min(G.cps, (last - s0) / 32 + 1);
#pragma unroll
        for (int f = 0; f < long_nf<MODE>(); ++f) {
          const T* tot = A.ws + A.slot_of(f, k.dir, 0, k.j, k.b, k.li) + static_cast<int64_t>(seg) * G.cps;
          T* car = A.ws + A.slot_of(f, k.dir, 1, k.j, k.b, k.li) + static_cast<int64_t>(seg) * G.cps;
          if (long_op<MODE>(f) == kVMax) carry_scan<T, kVMax>(tot, car, chunks, lane);
          else carry_scan<T, kVMin>(tot, car, chunks, lane);
        }
      }
      continue;
    }
    for (int32_t seg = blockIdx.x * 256 + threadIdx.x; seg < G.nseg; seg += gridDim.x * 256) {
      const int32_t s0 = G.o + (G.k_first + seg) * G.wa;
      const int32_t chunks = min(G.cps, (last - s0) / 32 + 1);  // those the folds reach
#pragma unroll
      for (int f = 0; f < long_nf<MODE>(); ++f) {
        const T* tot = A.ws + A.slot_of(f, k.dir, 0, k.j, k.b, k.li) + static_cast<int64_t>(seg) * G.cps;
        T* car = A.ws + A.slot_of(f, k.dir, 1, k.j, k.b, k.li) + static_cast<int64_t>(seg) * G.cps;
        if (long_op<MODE>(f) == kVMax) carry_run<T, kVMax>(tot, car, chunks);
        else if (long_op<MODE>(f) == kVMin) carry_run<T, kVMin>(tot, car, chunks);
        else carry_run<T, kVSum>(tot, car, chunks);
      }
    }
  }
}

// Step 3: every output from the two running folds and two carries: G at e = i + r of the line,
// H(i - r) = G at n - 1 - i + r of the reversed line (the workgroups of a grid row stride over the
// positions of one line).
template <class T, int MODE>
__global__ void __launch_bounds__(256) k_long_out(LongArgs<T> A) {
  __shared__ LongTask st;
  const LongPass& a = A.a;
  const int64_t tasks = static_cast<int64_t>(a.rungs) * A.B * a.lines;
  for (int64_t t = blockIdx.y; t < tasks; t += gridDim.y) {
    if (threadIdx.x == 0) long_task(A, t, false, &st);
    __syncthreads();
    const LongTask k = st;
    __syncthreads();
    if (k.w < 1) continue;
    const LongGeo& G = k.G;
    T fac = T(1);
    if (a.scale) {
      double prod = 1.0;
      for (int ax = 0; ax < a.g.dim; ++ax) prod *= static_cast<double>(2 * min((k.w - 1) / 2, a.g.n[ax] - 1) + 1);
      fac = static_cast<T>(1.0 / prod);
    }
    for (int32_t i = blockIdx.x * 256 + threadIdx.x; i < a.n; i += gridDim.x * 256) {
      const auto fold = [&](int f, int op, int dir, int32_t e, int32_t at) -> T {
        const int32_t d = e - G.o;
        const int32_t ks = d >= 0 ? d / G.wa : -((-d + G.wa - 1) / G.wa);
        const int32_t qc = (e - (G.o + ks * G.wa)) >> 5;
        const T local = A.pre[A.pre_of(f, dir, k.j, k.b, k.li) + at];  // the fold from the chunk start
        if (qc == 0) return local;
        const int64_t c = static_cast<int64_t>(ks - G.k_first) * G.cps + qc;
        return long_apply(op, A.ws[A.slot_of(f, dir, 1, k.j, k.b, k.li) + c], local);
      };
#pragma unroll
      for (int f = 0; f < long_nf<MODE>(); ++f) {
        const int op = long_op<MODE>(f);
        const T gv = fold(f, op, 0, i + G.r, i);
        const T hv = fold(f, op, 1, a.n - 1 - i + G.r, a.n - 1 - i);
        T v;
        if (op == kVSum) {
          const T sum = hv + gv;
          const int32_t m = (i - G.r - G.o) % G.wa;
          v = m == 0 ? sum * T(0.5) : sum;  // (i - r - o) mod wa == 0: the window is one segment
          if (a.scale) v = v * fac;
        } else {
          v = long_apply(op, hv, gv);
        }
        A.out[f][k.j * A.out_stride + k.base + static_cast<int64_t>(i) * a.step] = v;
      }
    }
  }
}

template <class T>
void launch_long(hipStream_t s, const Geom& g, int axis, int nf, const int* ops, bool scale, const VPtrs<T>& p,
                 int32_t rungs, int32_t max_w) {
  LongPass a{};
  a.g = g;
  a.n = g.n[axis];
  a.nf = nf;
  a.ops[0] = ops[0];
  a.ops[1] = nf > 1 ? ops[1] : ops[0];
  a.scale = scale ? 1 : 0;
  a.rungs = rungs;
  a.step = g.st[axis];
  a.inner = axis == g.dim - 1 ? 1 : g.st[axis];
  a.lines = g.count / a.n;
  // chunks of a line and direction: the segments meeting [-r, n - 1 + r] (at most L / wa + 2 of
  // them, L = n + 2r) times ceil(wa / 32) chunks each, at most 1.1 L + 2
  const int32_t rmax = std::min((max_w - 1) / 2, a.n - 1);
  const int64_t L = static_cast<int64_t>(a.n) + 2 * rmax;
  a.cap = L + L / 8 + 64;
  // chunk totals and carries of both directions, per field, line, element and rung
  const int64_t per_dir = static_cast<int64_t>(rungs) * g.batch * a.lines * a.cap;
  const size_t fold_bytes = static_cast<size_t>(per_dir) * 2 * 2 * static_cast<size_t>(nf) * sizeof(T);
  // The outputs fold the input from its chunk starts while other threads write outputs: a pass
  // whose input and output share memory (every pass after a cube's first runs in place) reads a
  // copy of its input.
  const int64_t in_rungs = p.in_stride == 0 ? 1 : rungs;
  const int64_t elems = g.batch * g.count;
  const auto span = [&](const T* q, int64_t rs, int64_t stride) {
    const uintptr_t lo = reinterpret_cast<uintptr_t>(q);
    return std::make_pair(lo, lo + static_cast<uintptr_t>(((rs - 1) * stride + elems) * sizeof(T)));
  };
  bool alias = false;
  for (int f = 0; f < nf; ++f)
    for (int h = 0; h < nf; ++h) {
      const auto i = span(p.in[f], in_rungs, p.in_stride);
      const auto o = span(p.out[h], rungs, p.out_stride);
      alias = alias || (i.first < o.second && o.first < i.second);
    }
  const int64_t stage_elems = alias ? static_cast<int64_t>(nf) * in_rungs * elems : 0;
  // the running folds from each chunk start, at the positions the outputs read ([r, n - 1 + r] of
  // both directions), per field, direction, rung and element
  const int64_t elems_all = g.batch * g.count;
  const size_t pre_elems = static_cast<size_t>(nf) * 2 * static_cast<size_t>(rungs) * elems_all;
  const size_t need = fold_bytes + (pre_elems + static_cast<size_t>(stage_elems)) * sizeof(T);
  // one workspace per (device, stream) that only grows: launches on one stream are ordered
  static std::mutex mu;
  static std::map<std::pair<int, hipStream_t>, std::pair<void*, size_t>> cache;
  int dev = 0;
  EDA_HIP(hipGetDevice(&dev));
  T* ws = nullptr;
  {
    std::lock_guard<std::mutex> lock(mu);
    auto& slot = cache[std::make_pair(dev, s)];
    if (slot.second < need) {
      EDA_HIP(hipStreamSynchronize(s));
      if (slot.first) EDA_HIP(hipFree(slot.first));
      slot.first = nullptr;
      const hipError_t e = hipMalloc(&slot.first, need);
      if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        slot.second = 0;
        throw OutOfDeviceMemory("eda: the device memory does not hold the long-line workspace");
      }
      EDA_HIP(e);
      slot.second = need;
    }
    ws = static_cast<T*>(slot.first);
  }
  const T* in0 = p.in[0];
  const T* in1 = p.in[1];
  int64_t in_stride = p.in_stride;
  if (alias) {
    T* stage = ws + fold_bytes / sizeof(T) + pre_elems;
    const T* src0 = in0;
    const T* src1 = in1;
    const int64_t src_stride = p.in_stride;
    flat(s, stage_elems, [=] __device__(int64_t it) {
      const int64_t k = it % elems;
      const int64_t q = it / elems;
      const int64_t j = q % in_rungs;
      const T* src = q / in_rungs == 0 ? src0 : src1;
      stage[it] = src[j * src_stride + k];
    });
    in0 = stage;
    in1 = nf > 1 ? stage + in_rungs * elems : stage;
    in_stride = p.in_stride == 0 ? 0 : elems;
  }
  const int32_t* win = p.win;
  const int64_t B = g.batch;
  T* const pre = ws + fold_bytes / sizeof(T);
  LongArgs<T> A{};
  A.a = a;
  A.win = win;
  A.in[0] = in0;
  A.in[1] = in1;
  A.out[0] = p.out[0];
  A.out[1] = p.out[1];
  A.in_stride = in_stride;
  A.out_stride = p.out_stride;
  A.B = B;
  A.ws = ws;
  A.pre = pre;
  const int64_t tasks = static_cast<int64_t>(rungs) * B * a.lines * 2;
  const unsigned gy = static_cast<unsigned>(std::min<int64_t>(tasks, 65535));
  // 1: the chunk folds and totals (about 4096 workgroups in all)
  constexpr int kW = fold_waves<T>();
  const int64_t gx1 = std::max<int64_t>(1, std::min<int64_t>((a.cap + 64 * kW - 1) / (64 * kW), (4096 + tasks - 1) / tasks));
  // 2: the carries (about 1024 workgroups)
  const int64_t segcap = L / 33 + 3;
  const int64_t gx2 = std::max<int64_t>(1, std::min<int64_t>((segcap + 255) / 256, (1024 + tasks - 1) / tasks));
  // 3: the outputs (about 4096 workgroups)
  const int64_t tasks3 = tasks / 2;
  const unsigned gy3 = static_cast<unsigned>(std::min<int64_t>(tasks3, 65535));
  const int64_t gx3 = std::max<int64_t>(1, std::min<int64_t>((a.n + 255) / 256, (4096 + tasks3 - 1) / tasks3));
  const int mode = nf == 2 ? kLongMaxMin : (ops[0] == kVMax ? kLongMax : (ops[0] == kVMin ? kLongMin : kLongSum));
  const auto go = [&](auto mode_c) {
    constexpr int M = decltype(mode_c)::value;
    EDA_GO((k_long_fold<T, M>), dim3(static_cast<unsigned>(gx1), gy), dim3(64 * kW), 0, s, A);
    EDA_GO((k_long_carry<T, M>), dim3(static_cast<unsigned>(gx2), gy), dim3(256), 0, s, A);
    EDA_GO((k_long_out<T, M>), dim3(static_cast<unsigned>(gx3), gy3), dim3(256), 0, s, A);
  };
  if (mode == kLongMax) go(std::integral_constant<int, kLongMax>{});
  else if (mode == kLongMin) go(std::integral_constant<int, kLongMin>{});
  else if (mode == kLongSum) go(std::integral_constant<int, kLongSum>{});
  else go(std::integral_constant<int, kLongMaxMin>{});
}

// The (rung, element) items of a launch whose clamped radius on lines of n samples lies in
// [rlo, rhi), from the host mirror of the window array; -1 when it is not mirrored.
int64_t count_radius(const int32_t* win, int64_t items, int32_t n, int32_t rlo, int32_t rhi) {
  const int32_t* h = static_cast<const int32_t*>(mirror_host(win, static_cast<size_t>(items) * sizeof(int32_t)));
  if (h == nullptr) return -1;
  int64_t c = 0;
  for (int64_t k = 0; k < items; ++k) {
    const int32_t rw = std::min((h[k] - 1) / 2, n - 1);
    c += h[k] >= 1 && rw >= rlo && rw < rhi ? 1 : 0;
  }
  return c;
}

// One pass on the staged kernels (a strided axis, a long line or a box sum along the contiguous
// axis); rmin = 16: the elements with shorter windows are served by the streaming pass and skipped.
template <class T>
void enqueue_staged(hipStream_t s, const Geom& g, int axis, int nf, bool same_input, const int* ops, bool scale,
                    const VPtrs<T>& p, int32_t max_w, int32_t rungs, int32_t rmin, bool timed) {
  const bool sum = ops[0] == kVSum;
  VPass a{};
  a.g = g;
  a.n = g.n[axis];
  a.step = g.st[axis];
  a.inner = axis == g.dim - 1 ? 1 : g.st[axis];
  a.outer = g.count / (static_cast<int64_t>(a.n) * a.inner);
  a.nf = nf;
  a.ops[0] = ops[0];
  a.ops[1] = nf > 1 ? ops[1] : ops[0];
  a.scale = scale ? 1 : 0;
  a.rmin = rmin;
  // the tile of a shape: lines, rows, workgroups, the 16-byte access test
  const auto shaped = [&](const Shape& sh) {
    VPass v = a;
    v.P = sh.P;
    v.lpad = sh.lpad;
    v.span = sh.span;
    v.flat = sh.flat;
    v.wrows = sh.wrows;
    v.epad = sh.epad;
    v.lead = sh.lead;
    if (v.inner == 1) {
      v.groups = (v.outer + v.P - 1) / v.P;
    } else {
      const int64_t gpo = (v.inner + v.P - 1) / v.P;
      v.groups = v.outer * gpo;
    }
    const int64_t wgs = g.batch * v.groups;
    if (wgs > 0x7fffffff) throw std::invalid_argument("eda: too many workgroups for one pass");
    v.xcd = xcd_chunk(wgs);
    const auto al16 = [](const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; };
    constexpr int64_t VW = 16 / static_cast<int64_t>(sizeof(T));
    // strided axis: P and the column count multiples of VW; contiguous axis: n a multiple of VW
    const bool shape_ok = v.inner > 1 ? v.P % VW == 0 && v.inner % VW == 0 : v.n % VW == 0;
    v.vec = shape_ok && al16(p.in[0]) && al16(p.in[1]) && al16(p.out[0]) && al16(p.out[1]) &&
                    (p.in_stride % VW) == 0 && (p.out_stride % VW) == 0
                ? 1
                : 0;
    return v;
  };
  hipEvent_t t0 = nullptr;
  int tag = kTagStrided;
  if (timed) timer_note(s, &t0);
  bool done = false;
  if (!pass_fits(a, sizeof(T), sum, max_w)) {
    launch_long<T>(s, g, axis, nf, ops, scale, p, rungs, max_w);
    EDA_LAUNCH("k_long");
    done = true;
  }
  if (!done && sum && EDA_VHGW_PIPE) {
    // the pipelined kernel, one launch per row layout: windows up to 32 samples on plain rows
    // (radius <= 15), longer ones on padded rows
    const bool shorter = rmin < 16 && count_radius(p.win, g.batch * rungs, a.n, rmin, 16) != 0;
    const bool longer = (max_w - 1) / 2 >= 16 && a.n > 16 &&
                        count_radius(p.win, g.batch * rungs, a.n, std::max(16, rmin), a.n) != 0;
    const bool pipe_fits = (!shorter || pass_fits(a, sizeof(T), true, std::min(max_w, 31), 1)) &&
                           (!longer || pass_fits(a, sizeof(T), true, max_w, 2));
    Shape s1{};
    VPass a1{};
    int pf1 = 4;
    if (pipe_fits && shorter) {
      s1 = pass_shape(a, sizeof(T), true, std::min(max_w, 31), 1);
      a1 = shaped(s1);
      pf1 = box_reads<T>(a1, 64 * s1.W);
    }
    Shape s2{};
    VPass a2{};
    int pf2 = 4;
    if (pipe_fits && longer) {
      s2 = pass_shape(a, sizeof(T), true, max_w, 2);
      a2 = shaped(s2);
      pf2 = box_reads<T>(a2, 64 * s2.W);
    }
    if (pipe_fits && pf1 && pf2 && (shorter || longer)) {
      const auto items = [&](bool longer_ones) {
        const int64_t c = count_radius(p.win, g.batch * rungs, a.n, longer_ones ? std::max(16, rmin) : rmin,
                                       longer_ones ? a.n : 16);
        return static_cast<double>(c < 0 ? g.batch * rungs : c);
      };
      if (shorter) {
        launch_box<T, false>(pf1, s, a1, p, rungs, s1.W, s1.lds);
        if (timed) {
          timer_done(s, t0, 2.0 * items(false) * g.count * sizeof(T), kTagBoxShort);
          t0 = nullptr;
        }
      }
      if (longer) {
        if (timed && t0 == nullptr) timer_note(s, &t0);
        launch_box<T, true>(pf2, s, a2, p, rungs, s2.W, s2.lds);
        if (timed) {
          timer_done(s, t0, 2.0 * items(true) * g.count * sizeof(T), kTagBoxLong);
          t0 = nullptr;
        }
      }
      EDA_LAUNCH("k_box");
      done = true;
    } else if (!shorter && !longer) {
      done = true;  // every element streamed
    }
  }
  if (!done) {
    const Shape sh = pass_shape(a, sizeof(T), sum, max_w);
    const VPass v = shaped(sh);
    const int64_t wgs = g.batch * v.groups;
    const dim3 grid(static_cast<unsigned>(wgs), static_cast<unsigned>(rungs));
    const dim3 block(static_cast<unsigned>(64 * sh.W));
    if (sum) launch_kt<T, T>(1, grid, block, sh.lds, s, v, p);
    else launch_kt<T, okey_t<T>>(kt_for(v.n), grid, block, sh.lds, s, v, p);
    EDA_LAUNCH("k_vhgw");
  }
  if (timed && t0 != nullptr) {
    int64_t active = mirror_active(p.win, g.batch * rungs);
    if (active < 0) active = g.batch * rungs;
    if (rmin > 0) {
      const int64_t c = count_radius(p.win, g.batch * rungs, a.n, rmin, a.n);
      if (c >= 0) active = c;
    }
    const int fields = (nf == 2 && !same_input ? 2 : 1) + nf;
    timer_done(s, t0, static_cast<double>(fields) * active * g.count * sizeof(T), tag);
  }
}

template <class T>
void enqueue_vhgw(hipStream_t s, const Geom& g, int axis, int nf, bool same_input, const int* ops, bool scale,
                  const T* const* in, T* const* out, const int32_t* win, int32_t max_w, int32_t rungs,
                  int64_t in_stride, int64_t out_stride, T* pairs = nullptr, int32_t rmin0 = 0) {
  if (max_w < 1 || rungs < 1) return;
  if (rungs > 65535) throw std::invalid_argument("eda: too many rungs for one launch");
  VPtrs<T> p{};
  p.in[0] = in[0];
  p.in[1] = nf > 1 ? in[1] : in[0];
  p.out[0] = out[0];
  p.out[1] = nf > 1 ? out[1] : out[0];
  p.win = win;
  p.in_stride = in_stride;
  p.out_stride = out_stride;
  const bool sum = ops[0] == kVSum;
  const bool il = pairs != nullptr && nf == 2 && !sum && rows_qualify(g, axis, sum);  // the row pass writes pairs
  if (il) {
    p.out[0] = pairs;
    p.out[1] = pairs;
    p.out_stride = 2 * out_stride;
  }
  const bool timed = timer_on();
  if (rows_qualify(g, axis, sum)) {
    // elements below rmin0 are the fused cube's; a launch with none of the others is not made
    const int64_t above = rmin0 > 0 ? count_radius(win, g.batch * rungs, g.n[axis], rmin0, g.n[axis]) : -1;
    if (above == 0) return;
    hipEvent_t t0 = nullptr;
    if (timed) timer_note(s, &t0);
    launch_rows<T>(s, g, axis, nf, same_input, ops, p, rungs, il, rmin0);
    EDA_LAUNCH("k_vhgw_rows");
    if (timed) {
      int64_t active = above >= 0 ? above : mirror_active(win, g.batch * rungs);
      if (active < 0) active = g.batch * rungs;
      const int fields = (nf == 2 && !same_input ? 2 : 1) + nf;
      timer_done(s, t0, static_cast<double>(fields) * active * g.count * sizeof(T), kTagRows);
    }
    return;
  }
  if (stream_qualifies<T>(g, axis, nf, same_input, ops, p)) {
    VPass a{};
    a.g = g;
    a.n = g.n[axis];
    a.step = g.st[axis];
    a.inner = g.st[axis];
    a.outer = g.count / (static_cast<int64_t>(a.n) * a.inner);
    a.nf = nf;
    // the staged kernels must hold the launch's longer windows (else the whole pass stays with them)
    if (max_w <= kStreamBigL || pass_fits(a, sizeof(T), sum, max_w)) {
      // elements with windows up to 31 samples stream in the short walk, up to 63 in the long one;
      // the launch's longer windows (if any) run the staged kernels, which skip the streamed elements
      const auto stream_class = [&](bool big) {
        hipEvent_t t0 = nullptr;
        if (timed) timer_note(s, &t0);
        launch_stream<T>(s, g, axis, nf, ops, scale, p, rungs, std::min(max_w, big ? kStreamBigL : kStreamMaxL), big,
                         rmin0);
        EDA_LAUNCH("k_vhgw_stream");
        if (timed) {
          const int64_t c = count_radius(win, g.batch * rungs, g.n[axis], big ? 16 : rmin0, big ? 32 : 16);
          const int fields = (nf == 2 && !same_input ? 2 : 1) + nf;
          timer_done(s, t0, static_cast<double>(fields) * (c < 0 ? g.batch * rungs : c) * g.count * sizeof(T),
                     kTagStream);
        }
      };
      // a class none of the launch's elements falls in is not launched (when the window array has
      // a host copy; else every class is)
      const int64_t items = g.batch * rungs;
      if (rmin0 < 16 && count_radius(win, items, g.n[axis], rmin0, 16) != 0) stream_class(false);
      if (max_w > kStreamMaxL && count_radius(win, items, g.n[axis], 16, 32) != 0) stream_class(true);
      if (max_w > kStreamBigL && count_radius(win, items, g.n[axis], 32, g.n[axis]) != 0)
        enqueue_staged<T>(s, g, axis, nf, same_input, ops, scale, p, max_w, rungs, 32, timed);
      return;
    }
  }
  enqueue_staged<T>(s, g, axis, nf, same_input, ops, scale, p, max_w, rungs, rmin0, timed);
}

// The cube (reference cpp/src/method/common.cpp:65-136): box runs axis 0 first, the last pass
// scaled; max / min run the last axis first; maxmin reads the input for both outputs. With
// rungs > 1 every pass runs all rungs in one launch: rung j filters in with the windows win + j B
// into out0 + j stride (out1 + j stride).
// A field per (device, stream) that only grows (launches on one stream are ordered): the first pass
// of an in-place cube writes it, so that the streaming pass of axis 0 does not run in place.
template <class T>
T* cube_scratch(hipStream_t s, int64_t elems) {
  static std::mutex mu;
  static std::map<std::pair<int, hipStream_t>, std::pair<void*, size_t>> cache;
  const size_t need = static_cast<size_t>(elems) * sizeof(T);
  int dev = 0;
  EDA_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  auto& slot = cache[std::make_pair(dev, s)];
  if (slot.second < need) {
    EDA_HIP(hipStreamSynchronize(s));
    if (slot.first) EDA_HIP(hipFree(slot.first));
    slot.first = nullptr;
    slot.second = 0;
    const hipError_t e = hipMalloc(&slot.first, need);
    if (e == hipErrorOutOfMemory) {
      (void)hipGetLastError();
      slot.first = nullptr;
      return nullptr;  // the cube then runs in place on the staged kernels
    }
    EDA_HIP(e);
    slot.second = need;
  }
  return static_cast<T*>(slot.first);
}

// The fields the passes of a max, min or box cube write, so that no pass reads the field it writes
// (the streaming passes need that): the last pass writes out0, the others alternate with a scratch
// field (the caller's, of `rungs` fields; else, for one rung, the stream's own). An in-place 3-D
// cube keeps its last pass in place. Without a scratch every pass writes out0.
template <class T>
void cube_targets(hipStream_t s, const Geom& g, const T* in, T* out0, int32_t rungs, T* scratch, T** dst) {
  const int D = g.dim;
  for (int a = 0; a < D; ++a) dst[a] = out0;
  if (D < 2) return;
  T* scr = scratch;
  if (scr == nullptr && rungs == 1 && static_cast<double>(g.count) * sizeof(T) < 4294967295.0)
    scr = cube_scratch<T>(s, g.batch * g.count);
  if (scr == nullptr) return;
  if (D == 2) {
    dst[0] = scr;
  } else if (in != out0) {
    dst[1] = scr;
  } else {
    dst[0] = scr;
  }
}

// The fused 2-D cube for the launch's elements with windows up to 31 samples (two classes, each
// launched when one of its elements runs). Returns the radius from which the two-pass path takes
// the elements (0: the fused cube does not serve the launch; 16, or 8 when the longer class does
// not fit the LDS: it served the shorter windows), or -1 when no element is left to the two-pass
// path.
template <class T>
int32_t enqueue_cube2(hipStream_t s, const Geom& g, int mode, bool pair, bool scale, const VPtrs<T>& p,
                      int32_t rungs, int32_t max_w) {
  if (!cube2_qualifies<T>(g, mode, pair, p, rungs)) return 0;
  const int64_t items = g.batch * rungs;
  const bool timed = timer_on();
  const int32_t n1 = g.n[1];  // radii below 16 are the same on both axes (n0 >= 128, n1 >= 64)
  // max + min: windows up to 7 samples (its three staged regions of 8 rows leave room for two
  // workgroups per CU); the others: up to 15, and 17 .. 31 where the LDS holds 2 x 31 rows
  const bool mid_ok = mode != 2 && cube2_lds<T>(n1, mode, true) <= 160 * 1024;
  const int32_t small_hi = mode == 2 ? 4 : 8;
  for (int cls = 0; cls < 2; ++cls) {
    if (cls == 1 && (max_w < 17 || !mid_ok)) break;
    const int64_t c = count_radius(p.win, items, n1, cls ? 8 : 0, cls ? 16 : small_hi);
    if (c == 0) continue;
    hipEvent_t t0 = nullptr;
    if (timed) timer_note(s, &t0);
    launch_cube2<T>(s, g, mode, pair, scale, p, rungs, cls == 1);
    EDA_LAUNCH("k_cube2");
    if (timed) {
      const int fields = 1 + (mode == 2 ? 2 : 1);
      timer_done(s, t0, static_cast<double>(fields) * (c < 0 ? items : c) * g.count * sizeof(T), kTagCube2);
    }
  }
  const int32_t from = mid_ok ? 16 : small_hi;
  if (max_w < 2 * from + 1) return -1;
  return count_radius(p.win, items, n1, from, n1) == 0 ? -1 : from;
}

template <class T>
bool enqueue_cube_rungs(hipStream_t s, const Geom& g, CubeOp op, const T* in, T* out0, T* out1, const int32_t* win,
                        int32_t rungs, int64_t stride, int32_t max_w, T* pairs, T* scratch) {
  if (max_w < 1) return false;
  const int D = g.dim;
  // The max and min side by side: the streaming row pass writes them, so it runs last (axes 0 ..
  // D - 1; max and min give the same bits in any order of the axes).
  const bool il = pairs != nullptr && op == CubeOp::kMaxMin && D >= 2 && rows_qualify(g, D - 1, false);
  // 2-D: the fused cube serves the windows up to 31 samples, the passes below the others
  int32_t rmin0 = 0;
  if (D == 2) {
    VPtrs<T> p{};
    p.in[0] = in;
    p.in[1] = in;
    p.out[0] = il ? pairs : out0;
    p.out[1] = il ? pairs : (out1 != nullptr ? out1 : out0);
    p.win = win;
    p.in_stride = 0;
    p.out_stride = il ? 2 * stride : stride;
    const int mode = op == CubeOp::kBox ? 3 : op == CubeOp::kMaxMin ? 2 : op == CubeOp::kMin ? 1 : 0;
    if (mode != 2 || EDA_CUBE2_MAXMIN) {
      rmin0 = enqueue_cube2<T>(s, g, mode, il, op == CubeOp::kBox, p, rungs, max_w);
      if (rmin0 < 0) return il;
    }
  }
  if (op == CubeOp::kBox) {
    T* dst[3];
    cube_targets<T>(s, g, in, out0, rungs, scratch, dst);
    for (int a = 0; a < D; ++a) {
      const int ops[1] = {kVSum};
      const T* src[1] = {a == 0 ? in : dst[a - 1]};
      T* out[1] = {dst[a]};
      enqueue_vhgw<T>(s, g, a, 1, true, ops, a == D - 1, src, out, win, max_w, rungs, a == 0 ? 0 : stride, stride,
                      nullptr, rmin0);
    }
    return false;
  }
  T* dst[3];
  if (op != CubeOp::kMaxMin) cube_targets<T>(s, g, in, out0, rungs, scratch, dst);
  for (int k = 0; k < D; ++k) {
    // axis 0 first: its pass reads the input and writes another field, as the streaming pass of a
    // strided axis needs
    const int a = k;
    const bool first = k == 0;
    const int64_t ins = first ? 0 : stride;
    if (op == CubeOp::kMaxMin) {
      const int ops[2] = {kVMax, kVMin};
      const T* src[2] = {first ? in : out0, first ? in : out1};
      T* out[2] = {out0, out1};
      enqueue_vhgw<T>(s, g, a, 2, first, ops, false, src, out, win, max_w, rungs, ins, stride,
                      il && k == D - 1 ? pairs : nullptr, rmin0);
    } else {
      const int ops[1] = {op == CubeOp::kMax ? kVMax : kVMin};
      const T* src[1] = {first ? in : dst[a - 1]};
      T* out[1] = {dst[a]};
      enqueue_vhgw<T>(s, g, a, 1, true, ops, false, src, out, win, max_w, rungs, ins, stride, nullptr, rmin0);
    }
  }
  return il;
}

template <class T>
void enqueue_cube(hipStream_t s, const Geom& g, CubeOp op, const T* in, T* out0, T* out1, const int32_t* win,
                  int32_t max_w) {
  enqueue_cube_rungs<T>(s, g, op, in, out0, out1, win, 1, 0, max_w, nullptr, nullptr);
}

template <class T>
void enqueue_axis_pass(hipStream_t s, const Geom& g, int axis, bool min, const T* in, T* out, const int32_t* win,
                       int32_t max_w, int32_t rungs, int64_t in_stride, int64_t out_stride) {
  const int ops[1] = {min ? kVMin : kVMax};
  const T* src[1] = {in};
  T* dst[1] = {out};
  enqueue_vhgw<T>(s, g, axis, 1, true, ops, false, src, dst, win, max_w, rungs, in_stride, out_stride);
}
template void enqueue_axis_pass<float>(hipStream_t, const Geom&, int, bool, const float*, float*, const int32_t*,
                                       int32_t, int32_t, int64_t, int64_t);
template void enqueue_axis_pass<double>(hipStream_t, const Geom&, int, bool, const double*, double*, const int32_t*,
                                        int32_t, int32_t, int64_t, int64_t);

template void enqueue_cube<float>(hipStream_t, const Geom&, CubeOp, const float*, float*, float*, const int32_t*, int32_t);
template void enqueue_cube<double>(hipStream_t, const Geom&, CubeOp, const double*, double*, double*, const int32_t*,
                                   int32_t);
template bool enqueue_cube_rungs<float>(hipStream_t, const Geom&, CubeOp, const float*, float*, float*, const int32_t*,
                                        int32_t, int64_t, int32_t, float*, float*);
template bool enqueue_cube_rungs<double>(hipStream_t, const Geom&, CubeOp, const double*, double*, double*,
                                         const int32_t*, int32_t, int64_t, int32_t, double*, double*);

}  // namespace eda
