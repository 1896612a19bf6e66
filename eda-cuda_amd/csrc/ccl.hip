// SPDX-License-Identifier: Apache-2.0
// Face-connected components (scipy.ndimage.label's default structure; the domain border does not
// connect), by lock-free union-find (Playne and Hawick's scheme; reference
// cpp/src/method/stats_scale.cu:1-83): a link always hangs the larger root under the smaller one by
// atomicMin, so every parent is smaller than its child and a component's root is its smallest
// in-element index whatever the order of the atomics; a last pass writes every voxel's root.
//
// gfx950 mapping. Along the contiguous axis a run of mask voxels is one component piece whose
// smallest index is its first voxel: one wave per line labels every voxel with its run's first
// voxel by a max-scan of the run starts (so runs need no links). Across the other axes a voxel v
// links with v + st only where the two runs begin to touch: when v - 1 and v - 1 + st are not both
// in the mask (otherwise that earlier pair already joins the same two runs).
#include "common.hpp"
#include "kernels.hpp"

namespace eda {
namespace {

__device__ __forceinline__ int32_t find_root(const int32_t* L, int32_t x) {
  const volatile int32_t* V = L;
  int32_t p = V[x];
  while (p != x) {
    x = p;
    p = V[x];
  }
  return x;
}

__device__ __forceinline__ void unite(int32_t* L, int32_t a, int32_t c) {
  for (;;) {
    a = find_root(L, a);
    c = find_root(L, c);
    if (a == c) return;
    if (a > c) {
      const int32_t t = a;
      a = c;
      c = t;
    }
    const int32_t old = atomicMin(L + c, a);
    if (old == c) return;
    c = old;
  }
}

// Inclusive max-scan over the 64 lanes of a wave.
__device__ __forceinline__ int32_t wave_max_scan(int32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int32_t u = __shfl_up(v, d, 64);
    if (lane >= d) v = u > v ? u : v;
  }
  return v;
}

constexpr int kRun = 8;                     // samples per lane per tile
constexpr int32_t kPiece = 64 * kRun * 8;   // samples per wave: lines longer than this split
constexpr int32_t kPending = INT32_MIN;     // the label of a run that began in an earlier piece

// One wave per piece of kPiece samples of a line of the contiguous (last) axis: labels = the index
// of the run's first voxel for mask voxels, -1 elsewhere; kPending for the leading voxels of a piece
// whose run began in an earlier piece (k_ccl_pending resolves them).
__global__ void __launch_bounds__(kThreads) k_ccl_runs(Geom g, const uint8_t* __restrict__ mask, uint8_t bits,
                                                       int32_t* __restrict__ labels, const int32_t* __restrict__ on,
                                                       int32_t pieces) {
  const int64_t b = blockIdx.y;
  if (on != nullptr && on[b] == 0) return;
  const int32_t n = g.n[g.dim - 1];
  const int64_t lines = g.count / n;
  const int lane = threadIdx.x & 63;
  const int64_t wv = static_cast<int64_t>(blockIdx.x) * (kThreads / 64) + (threadIdx.x >> 6);
  if (wv >= lines * pieces) return;  // uniform over the wave
  const int64_t line = wv / pieces;
  const int32_t p0 = static_cast<int32_t>(wv - line * pieces) * kPiece;
  const int32_t p1 = min(n, p0 + kPiece);
  const uint8_t* m = mask + b * g.count + line * n;
  int32_t* out = labels + b * g.count;
  const int32_t first = static_cast<int32_t>(line * n);  // the in-element index of the line's sample 0
  // the last run start of the earlier tiles (-2: the piece's leading run began before it)
  int32_t carry = p0 > 0 && (m[p0 - 1] & bits) != 0 ? -2 : -1;
  for (int32_t t0 = p0; t0 < p1; t0 += 64 * kRun) {
    const int32_t i0 = t0 + lane * kRun;
    bool prev = i0 > 0 && i0 - 1 < n && (m[i0 - 1] & bits) != 0;
    bool in[kRun];
    int32_t st[kRun];
    int32_t run = -1;
#pragma unroll
    for (int k = 0; k < kRun; ++k) {
      const int32_t i = i0 + k;
      in[k] = i < p1 && (m[i] & bits) != 0;
      if (in[k] && !prev) run = i;
      st[k] = run;
      prev = in[k];
    }
    const int32_t incl = wave_max_scan(run);
    int32_t excl = __shfl_up(incl, 1, 64);
    if (lane == 0) excl = -1;
    excl = excl > carry ? excl : carry;
#pragma unroll
    for (int k = 0; k < kRun; ++k) {
      const int32_t i = i0 + k;
      if (i < p1) {
        const int32_t s0 = st[k] >= 0 ? st[k] : excl;
        out[first + i] = !in[k] ? -1 : (s0 >= 0 ? first + s0 : kPending);
      }
    }
    const int32_t last = __shfl(incl, 63, 64);
    carry = last > carry ? last : carry;
  }
}

// One wave per piece after the first: its leading run, when it began in an earlier piece, takes the
// label of that run's start, found at the last voxel of the nearest earlier piece whose label is
// final (a piece whose last voxel is pending lies inside the run). Another wave may finalize such a
// voxel meanwhile: it then holds the same label.
__global__ void __launch_bounds__(kThreads) k_ccl_pending(Geom g, int32_t* __restrict__ labels,
                                                          const int32_t* __restrict__ on, int32_t pieces) {
  const int64_t b = blockIdx.y;
  if (on != nullptr && on[b] == 0) return;
  const int32_t n = g.n[g.dim - 1];
  const int64_t lines = g.count / n;
  const int lane = threadIdx.x & 63;
  const int64_t wv = static_cast<int64_t>(blockIdx.x) * (kThreads / 64) + (threadIdx.x >> 6);
  if (wv >= lines * pieces) return;
  const int64_t line = wv / pieces;
  const int32_t pc = static_cast<int32_t>(wv - line * pieces);
  if (pc == 0) return;
  const int32_t p0 = pc * kPiece;
  const int32_t p1 = min(n, p0 + kPiece);
  volatile int32_t* L = labels + b * g.count + line * n;
  if (L[p0] != kPending) return;
  int32_t lab = kPending;
  for (int32_t q = p0 - 1; q >= 0 && lab == kPending; q -= kPiece) lab = L[q];
  for (int32_t t0 = p0; t0 < p1; t0 += 64) {
    const int32_t i = t0 + lane;
    const bool pend = i < p1 && L[i] == kPending;
    if (pend) L[i] = lab;
    if (__ballot(pend) != ~0ULL) break;  // the leading run ended in this block
  }
}

}  // namespace

void enqueue_ccl(hipStream_t s, const Geom& g, const uint8_t* mask, uint8_t bits, int32_t* labels, const int32_t* on) {
  const int32_t nl = g.n[g.dim - 1];
  const int64_t lines = g.count / nl;
  const int32_t pieces = (nl + kPiece - 1) / kPiece;
  const int64_t gx = (lines * pieces + kThreads / 64 - 1) / (kThreads / 64);
  EDA_GO(k_ccl_runs, dim3(static_cast<unsigned>(gx), static_cast<unsigned>(g.batch)), dim3(kThreads), 0, s,
                     g, mask, bits, labels, on, pieces);
  EDA_LAUNCH("k_ccl_runs");
  if (pieces > 1) {
    EDA_GO(k_ccl_pending, dim3(static_cast<unsigned>(gx), static_cast<unsigned>(g.batch)), dim3(kThreads), 0, s,
                       g, labels, on, pieces);
    EDA_LAUNCH("k_ccl_pending");
  }
  if (g.dim == 1) return;  // a line's runs are its components
  each(s, g, on, [=] __device__(int64_t b, int64_t i64) {
    int32_t* L = labels + b * g.count;
    const int32_t v = static_cast<int32_t>(i64);
    if (L[v] < 0) return;
    const uint32_t uv = static_cast<uint32_t>(v);
    const bool left = uv % static_cast<uint32_t>(nl) != 0;  // v - 1 lies on v's line
    const bool left_in = left && L[v - 1] >= 0;
    for (int a = 0; a < g.dim - 1; ++a) {
      const uint32_t st = static_cast<uint32_t>(g.st[a]);
      const uint32_t na = static_cast<uint32_t>(g.n[a]);
      const uint32_t coord = a == 0 ? uv / st : (uv / st) % na;
      if (coord + 1 >= na) continue;
      const int32_t u = v + static_cast<int32_t>(st);
      if (L[u] < 0) continue;
      if (left_in && L[u - 1] >= 0) continue;  // the pair (v - 1, u - 1) joins the same runs
      unite(L, v, u);
    }
  });
  each(s, g, on, [=] __device__(int64_t b, int64_t i64) {
    int32_t* L = labels + b * g.count;
    const int32_t v = static_cast<int32_t>(i64);
    if (L[v] < 0) return;
    const int32_t root = find_root(L, v);
    reinterpret_cast<volatile int32_t*>(L)[v] = root;
  });
}

}  // namespace eda
