// SPDX-License-Identifier: Apache-2.0
// The variable structuring element of a mode and its operators (reference
// cpp/src/method/element.hpp:1-29, element.cuh:53-75, element.cu:166-303, element_op.cuh:1-263):
//   d = L ((1 - c) + c / q) the cube side, e = 0.5 max(L (1 + (span - 1) c) - d, 0) the segment
//   half length, o = e n the segment half vector (D >= 2); the hat weights in log d over the rung
//   ladder (lo = the lower rung, t = the weight of rung lo + 1); per rung j, the cube of side
//   rung_w[j] and a gather: acc = f(x), or over the 5 points x + (i / 2) o, i = -2..2 read by
//   D-linear interpolation of the half-sample reflected periodic extension when the rung carries
//   segment points (pts bit j); max / min, or the left-to-right sum divided by 5; the voxel's first
//   rung with a positive weight stores weight * acc, the second adds it. Point coordinates are
//   formed in double for both dtypes.
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <cmath>

namespace eda {
namespace {

template <class T>
struct DE {
  T d, e;
};
template <class T>
__device__ __forceinline__ DE<T> element_de(T L, T c, T q, T span_m1) {
  const T f = (T(1) - c) + c / q;
  DE<T> r;
  r.d = L * f;
  const T len = L * (T(1) + span_m1 * c);
  const T diff = len - r.d;
  r.e = T(0.5) * (diff > T(0) ? diff : T(0));
  return r;
}

// Workgroups per element: enough to fill the device for a batch, few enough that the per-element
// integer atomics of the build and the weights (one per workgroup) do not contend.
dim3 elem_grid(const Geom& g, int64_t cap = 1024) {
  int64_t bx = (g.count + kThreads - 1) / kThreads;
  const int64_t want = std::max<int64_t>(16, (int64_t{8192} + g.batch - 1) / g.batch);
  bx = std::min(bx, std::min(cap, want));
  bx = bx < 1 ? 1 : bx;
  return dim3(static_cast<unsigned>(bx), static_cast<unsigned>(g.batch));
}

template <class T>
__global__ void __launch_bounds__(kThreads) k_build(Geom g, T q, T span_m1, const T* __restrict__ L,
                                                    const T* __restrict__ nvec, const T* __restrict__ coh,
                                                    T* __restrict__ o, const int32_t* __restrict__ on,
                                                    uint64_t* __restrict__ dminmax) {
  __shared__ uint64_t smin[kThreads / 64], smax[kThreads / 64];
  const int64_t b = blockIdx.y;
  if (on != nullptr && on[b] == 0) return;
  const int64_t F = g.batch * g.count;
  uint64_t tmin = 0x7ff0000000000000ULL, tmax = 0;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kThreads) + threadIdx.x; i < g.count;
       i += static_cast<int64_t>(gridDim.x) * kThreads) {
    const int64_t at = b * g.count + i;
    const T c = coh != nullptr ? coh[at] : T(0);
    const DE<T> de = element_de<T>(L[at], c, q, span_m1);
    if (g.dim >= 2)
      for (int a = 0; a < g.dim; ++a) o[a * F + at] = de.e * nvec[a * F + at];
    const uint64_t bits = dbits(static_cast<double>(de.d));
    tmin = bits < tmin ? bits : tmin;
    tmax = bits > tmax ? bits : tmax;
  }
  tmin = wave_min_u64(tmin);
  tmax = wave_max_u64(tmax);
  if ((threadIdx.x & 63) == 0) {
    smin[threadIdx.x >> 6] = tmin;
    smax[threadIdx.x >> 6] = tmax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kThreads / 64; ++w) {
      tmin = smin[w] < tmin ? smin[w] : tmin;
      tmax = smax[w] > tmax ? smax[w] : tmax;
    }
    tmin = smin[0] < tmin ? smin[0] : tmin;
    tmax = smax[0] > tmax ? smax[0] : tmax;
    atomicMin(reinterpret_cast<unsigned long long*>(&dminmax[2 * b]), static_cast<unsigned long long>(tmin));
    atomicMax(reinterpret_cast<unsigned long long*>(&dminmax[2 * b + 1]), static_cast<unsigned long long>(tmax));
  }
}

template <class T>
__global__ void __launch_bounds__(kThreads) k_weights(Geom g, T q, T span_m1, const T* __restrict__ L,
                                                      const T* __restrict__ coh, const ElemDev* __restrict__ el,
                                                      uint8_t* __restrict__ lo, T* __restrict__ t,
                                                      const int32_t* __restrict__ on, uint32_t* __restrict__ bits_out) {
  __shared__ double s_lw[kMaxRungs];
  __shared__ uint32_t s_used, s_pts;
  const int64_t b = blockIdx.y;
  if (on != nullptr && on[b] == 0) return;
  const int K = el[b].n_rungs;
  // log(rung side) with the device log, the same function as the voxels' log d: a d clipped to a
  // rung gets a weight of exactly 0 or 1
  if (threadIdx.x < kMaxRungs) s_lw[threadIdx.x] = threadIdx.x < static_cast<unsigned>(K) ? log(static_cast<double>(el[b].rung_w[threadIdx.x])) : 0.0;
  if (threadIdx.x == 0) {
    s_used = 0u;
    s_pts = 0u;
  }
  __syncthreads();
  const double w_lo = static_cast<double>(el[b].rung_w[0]);
  const double w_hi = static_cast<double>(el[b].rung_w[K > 0 ? K - 1 : 0]);
  uint32_t used = 0u, pts = 0u;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kThreads) + threadIdx.x; i < g.count;
       i += static_cast<int64_t>(gridDim.x) * kThreads) {
    const int64_t at = b * g.count + i;
    const T c = coh != nullptr ? coh[at] : T(0);
    const DE<T> de = element_de<T>(L[at], c, q, span_m1);
    int idx = 0;
    T tt = T(0);
    if (K > 1) {
      double dc = static_cast<double>(de.d);
      dc = dc > w_lo ? dc : w_lo;
      dc = dc < w_hi ? dc : w_hi;
      const double lg = log(dc);
      int cnt = 0;
      for (int j = 0; j < K; ++j) cnt += s_lw[j] <= lg ? 1 : 0;
      idx = cnt - 1;
      idx = idx < 0 ? 0 : (idx > K - 2 ? K - 2 : idx);
      tt = static_cast<T>((lg - s_lw[idx]) / (s_lw[idx + 1] - s_lw[idx]));
    }
    lo[at] = static_cast<uint8_t>(idx);
    t[at] = tt;
    const T wl = T(1) - tt;
    uint32_t bits = 0u;
    if (wl > T(0)) bits |= 1u << idx;
    if (tt > T(0)) bits |= 1u << (idx + 1);
    used |= bits;
    if (de.e > T(0.25)) pts |= bits;
  }
  used = wave_or_u32(used);
  pts = wave_or_u32(pts);
  if ((threadIdx.x & 63) == 0) {
    if (used) atomicOr(&s_used, used);
    if (pts) atomicOr(&s_pts, pts);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s_used) atomicOr(&bits_out[2 * b], s_used);
    if (s_pts) atomicOr(&bits_out[2 * b + 1], s_pts);
  }
}

// Index i of the half-sample reflected periodic extension of an axis of n samples, in 32 bits (an
// element holds at most 2^31 samples): i itself inside the axis, else through i mod 2n.
__device__ __forceinline__ int32_t reflect32(int32_t i, int32_t n) {
  if (static_cast<uint32_t>(i) < static_cast<uint32_t>(n)) return i;
  const int32_t p = 2 * n;
  int32_t m = i % p;
  m += m < 0 ? p : 0;
  return m < n ? m : p - 1 - m;
}

// Two and four consecutive values read by one load at the alignment of one value (two): the
// hardware serves 8- and 16-byte global loads at 4-byte alignment.
template <class T>
struct Wide;
template <>
struct Wide<float> {
  typedef float two __attribute__((ext_vector_type(2), aligned(4)));
  typedef float four __attribute__((ext_vector_type(4), aligned(8)));
};
template <>
struct Wide<double> {
  typedef double two __attribute__((ext_vector_type(2), aligned(8)));
  typedef double four __attribute__((ext_vector_type(4), aligned(16)));
};

// D-linear interpolation at y of the half-sample reflected periodic extension of a field of one
// element: the corner offsets and weights once (setup), then any field with that layout (at).
template <class T, int D>
struct Interp {
  int32_t off0[D], off1[D];
  T w0[D], w1[D];
  __device__ __forceinline__ void setup(const int32_t* n, const int32_t* st, const double* y) {
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const double fl = floor(y[a]);
      const double tt = y[a] - fl;
      // |y| stays far below 2^30 for finite fields; the clamp keeps any other input in range
      const int32_t i0 = static_cast<int32_t>(fmin(fmax(fl, -1.0e9), 1.0e9));
      const T t1 = static_cast<T>(tt);
      w1[a] = t1;
      w0[a] = T(1) - t1;
      off0[a] = reflect32(i0, n[a]) * st[a];
      off1[a] = reflect32(i0 + 1, n[a]) * st[a];
    }
  }
  // Corners 2k and 2k + 1 differ in the last axis only: where its two samples are adjacent (inside
  // the axis; the last axis is contiguous) one read covers both, else two. Every corner's term and
  // the sums keep the corner order.
  __device__ __forceinline__ bool adjacent() const { return off1[D - 1] == off0[D - 1] + 1; }
  // the offset of corner 2k's sample (the last axis at its lower sample)
  __device__ __forceinline__ int32_t pair_off(int c2) const {
    int32_t off = off0[D - 1];
#pragma unroll
    for (int a = 0; a < D - 1; ++a) off += ((c2 >> (D - 1 - a)) & 1) ? off1[a] : off0[a];
    return off;
  }
  __device__ __forceinline__ T weigh(int corner, T v) const {
#pragma unroll
    for (int a = 0; a < D; ++a) v = v * (((corner >> (D - 1 - a)) & 1) ? w1[a] : w0[a]);
    return v;
  }
  // two fields stored side by side (field k of sample v at f + 2 v + k), each interpolated as at()
  // does
  __device__ __forceinline__ void at2(const T* __restrict__ f, T& r0, T& r1) const {
    using Four = typename Wide<T>::four;
    using Two = typename Wide<T>::two;
    const bool adj = adjacent();
    const int32_t step = off1[D - 1] - off0[D - 1];
    T acc0 = T(0), acc1 = T(0);
#pragma unroll
    for (int c2 = 0; c2 < (1 << D); c2 += 2) {
      const int32_t off = pair_off(c2);
      Four q;
      if (adj) {
        q = *reinterpret_cast<const Four*>(f + 2 * static_cast<int64_t>(off));
      } else {
        const Two lo = *reinterpret_cast<const Two*>(f + 2 * static_cast<int64_t>(off));
        const Two hi = *reinterpret_cast<const Two*>(f + 2 * static_cast<int64_t>(off + step));
        q.x = lo.x;
        q.y = lo.y;
        q.z = hi.x;
        q.w = hi.y;
      }
      acc0 = acc0 + weigh(c2, q.x);
      acc1 = acc1 + weigh(c2, q.y);
      acc0 = acc0 + weigh(c2 + 1, q.z);
      acc1 = acc1 + weigh(c2 + 1, q.w);
    }
    r0 = acc0;
    r1 = acc1;
  }
  __device__ __forceinline__ T at(const T* __restrict__ f) const {
    using Two = typename Wide<T>::two;
    const bool adj = adjacent();
    const int32_t step = off1[D - 1] - off0[D - 1];
    T acc = T(0);
#pragma unroll
    for (int c2 = 0; c2 < (1 << D); c2 += 2) {
      const int32_t off = pair_off(c2);
      Two q;
      if (adj) {
        q = *reinterpret_cast<const Two*>(f + off);
      } else {
        q.x = f[off];
        q.y = f[off + step];
      }
      acc = acc + weigh(c2, q.x);
      acc = acc + weigh(c2 + 1, q.y);
    }
    return acc;
  }
};

// The gather of every rung in one pass (OP 0 max, 2 max + min, 3 box): each voxel reads its lower
// rung lo, its weight t and (when one of its two rungs carries segment points) o once, and the cubes
// of its two rungs from the rung fields r0 (r1) + j F. The arithmetic is the per-rung one in rung
// order: the lower rung's weighted value stores (wl = 1 - t > 0), the upper rung's adds (t > 0).
// One rung's value at voxel i: f0[i] (f1[i]) folded with the 4 segment points when pts; the points
// in the order (1/2 o, -1/2 o, o, -o).
template <class T, int D, int OP>
__device__ __forceinline__ void rung_acc(int32_t i, const T* __restrict__ f0, const T* __restrict__ f1, bool pts,
                                         const Interp<T, D> (&ip)[4], T* a0, T* a1) {
  T acc0 = f0[i];
  T acc1 = OP == 2 ? f1[i] : T(0);
  if (pts) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if constexpr (OP == 2) {
        acc0 = max_c(acc0, ip[q].at(f0));
        acc1 = min_c(acc1, ip[q].at(f1));
      } else if constexpr (OP == 0) {
        acc0 = max_c(acc0, ip[q].at(f0));
      } else {
        acc0 = acc0 + ip[q].at(f0);
      }
    }
    if constexpr (OP == 3) acc0 = acc0 / T(5);
  }
  *a0 = acc0;
  *a1 = acc1;
}

// Max and min of one rung read side by side (pr + 2 v + k): the same arithmetic as rung_acc for op 2.
template <class T, int D>
__device__ __forceinline__ void rung_acc_pairs(int32_t i, const T* __restrict__ pr, bool pts,
                                               const Interp<T, D> (&ip)[4], T* a0, T* a1) {
  const typename Wide<T>::two q0 = *reinterpret_cast<const typename Wide<T>::two*>(pr + 2 * static_cast<int64_t>(i));
  T acc0 = q0.x;
  T acc1 = q0.y;
  if (pts) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      T m0, m1;
      ip[q].at2(pr, m0, m1);
      acc0 = max_c(acc0, m0);
      acc1 = min_c(acc1, m1);
    }
  }
  *a0 = acc0;
  *a1 = acc1;
}

// IL (op 2): the rung fields are the max and min side by side in r0 (rung j at r0 + 2 j F).
template <class T, int D, int OP, bool IL = false>
__global__ void __launch_bounds__(kThreads) k_gather_all(Geom g, const ElemDev* __restrict__ el,
                                                         const uint8_t* __restrict__ lo, const T* __restrict__ t,
                                                         const T* __restrict__ o, const T* __restrict__ r0,
                                                         const T* __restrict__ r1, T* __restrict__ out0,
                                                         T* __restrict__ out1, const int32_t* __restrict__ on,
                                                         GatherExtra<T> ex) {
  const int64_t b = blockIdx.y;
  if (on != nullptr && on[b] == 0) return;
  const uint32_t pts = D >= 2 ? el[b].pts : 0u;
  const int64_t base = b * g.count;
  const int64_t F = g.batch * g.count;
  const int32_t count = static_cast<int32_t>(g.count);  // at most 2^31 - 1 in this kernel's indices
  int32_t nn[D], st[D];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    nn[a] = g.n[a];
    st[a] = static_cast<int32_t>(g.st[a]);
  }
  for (int64_t i64 = blockIdx.x * static_cast<int64_t>(kThreads) + threadIdx.x; i64 < count;
       i64 += static_cast<int64_t>(gridDim.x) * kThreads) {
    const int32_t i = static_cast<int32_t>(i64);
    const int64_t at = base + i;
    const int l = lo[at];
    const T tt = t[at];
    const T wl = T(1) - tt;
    const bool use_lo = wl > T(0), use_up = tt > T(0);
    const bool pts_lo = use_lo && ((pts >> l) & 1u) != 0u;
    const bool pts_up = use_up && ((pts >> (l + 1)) & 1u) != 0u;
    // the 4 segment points of the voxel (shared by its two rungs), coordinates in double
    Interp<T, D> ip[4];
    if (pts_lo || pts_up) {
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
    T v0 = T(0), v1 = T(0);
    if (use_lo) {
      T a0, a1;
      if constexpr (IL) rung_acc_pairs<T, D>(i, r0 + 2 * (l * F + base), pts_lo, ip, &a0, &a1);
      else rung_acc<T, D, OP>(i, r0 + l * F + base, OP == 2 ? r1 + l * F + base : nullptr, pts_lo, ip, &a0, &a1);
      v0 = wl * a0;
      v1 = wl * a1;
    }
    if (use_up) {
      T a0, a1;
      if constexpr (IL)
        rung_acc_pairs<T, D>(i, r0 + 2 * ((l + 1) * F + base), pts_up, ip, &a0, &a1);
      else
        rung_acc<T, D, OP>(i, r0 + (l + 1) * F + base, OP == 2 ? r1 + (l + 1) * F + base : nullptr, pts_up, ip, &a0,
                           &a1);
      const T u0 = tt * a0, u1 = tt * a1;
      v0 = use_lo ? v0 + u0 : u0;
      v1 = use_lo ? v1 + u1 : u1;
    }
    out0[at] = v0;
    if constexpr (OP == 2) {
      out1[at] = v1;
      if (ex.mid != nullptr) ex.mid[at] = T(0.5) * (v0 + v1);
      if (ex.hdiff != nullptr) ex.hdiff[at] = T(0.5) * (v0 - v1);
      if (ex.pdiff != nullptr) {
        const T dv = v0 - v1;
        ex.pdiff[at] = dv >= T(0) ? dv : T(0);
      }
    }
    if constexpr (OP == 3) {
      if (ex.diff != nullptr) ex.diff[at] = ex.minuend[at] - v0;
    }
  }
}

template <class T, int OP>
void launch_gather_all(hipStream_t s, const Geom& g, const ElemDev* el, const uint8_t* lo, const T* t, const T* o,
                       const T* r0, const T* r1, T* out0, T* out1, const int32_t* on, const GatherExtra<T>& ex) {
  const dim3 grid = elem_grid(g);
  if (g.dim == 1)
    EDA_GO((k_gather_all<T, 1, OP>), grid, dim3(kThreads), 0, s, g, el, lo, t, o, r0, r1, out0, out1, on, ex);
  else if (g.dim == 2)
    EDA_GO((k_gather_all<T, 2, OP>), grid, dim3(kThreads), 0, s, g, el, lo, t, o, r0, r1, out0, out1, on, ex);
  else
    EDA_GO((k_gather_all<T, 3, OP>), grid, dim3(kThreads), 0, s, g, el, lo, t, o, r0, r1, out0, out1, on, ex);
  EDA_LAUNCH("k_gather_all");
}

// op 2 from the max and min side by side (D >= 2)
template <class T>
void launch_gather_pairs(hipStream_t s, const Geom& g, const ElemDev* el, const uint8_t* lo, const T* t, const T* o,
                         const T* pairs, T* out0, T* out1, const int32_t* on, const GatherExtra<T>& ex) {
  const dim3 grid = elem_grid(g);
  if (g.dim == 2)
    EDA_GO((k_gather_all<T, 2, 2, true>), grid, dim3(kThreads), 0, s, g, el, lo, t, o, pairs, nullptr, out0, out1, on,
           ex);
  else
    EDA_GO((k_gather_all<T, 3, 2, true>), grid, dim3(kThreads), 0, s, g, el, lo, t, o, pairs, nullptr, out0, out1, on,
           ex);
  EDA_LAUNCH("k_gather_all");
}

}  // namespace

template <class T>
void enqueue_element_build(hipStream_t s, const Geom& g, const MethodCtl& c, const T* L, const T* nvec, const T* coh,
                           T* o, const int32_t* on, uint64_t* dminmax) {
  EDA_GO(k_build<T>, elem_grid(g, 64), dim3(kThreads), 0, s, g, static_cast<T>(c.q),
                     static_cast<T>(c.span - 1.0), L, nvec, g.dim >= 2 ? coh : nullptr, o, on, dminmax);
  EDA_LAUNCH("k_build");
}

template <class T>
void enqueue_element_weights(hipStream_t s, const Geom& g, const MethodCtl& c, const T* L, const T* coh,
                             const ElemDev* el, uint8_t* lo, T* t, const int32_t* on, uint32_t* bits) {
  EDA_GO(k_weights<T>, elem_grid(g, 64), dim3(kThreads), 0, s, g, static_cast<T>(c.q),
                     static_cast<T>(c.span - 1.0), L, g.dim >= 2 ? coh : nullptr, el, lo, t, on, bits);
  EDA_LAUNCH("k_weights");
}

// The timer's algorithmic bytes of one gather: per voxel of a running element, lo (1 byte) and t
// read, the voxel's two bracketing rung fields (max and min: both) read, o (D fields) read where the
// element uses segment points, the outputs and extra outputs written (and the minuend read).
template <class T>
double gather_bytes(const Geom& g, int op, const ElemDev* el, const int32_t* on, const GatherExtra<T>& ex) {
  const ElemDev* he = static_cast<const ElemDev*>(mirror_host(el, static_cast<size_t>(g.batch) * sizeof(ElemDev)));
  const int32_t* ho =
      on != nullptr ? static_cast<const int32_t*>(mirror_host(on, static_cast<size_t>(g.batch) * sizeof(int32_t)))
                    : nullptr;
  const int nf = op == 2 ? 2 : 1;
  const int fields = 1 + 2 * nf + nf + (ex.mid ? 1 : 0) + (ex.hdiff ? 1 : 0) + (ex.pdiff ? 1 : 0) +
                     (ex.diff ? 2 : 0);
  double total = 0.0;
  for (int64_t b = 0; b < g.batch; ++b) {
    if (on != nullptr && ho != nullptr && ho[b] == 0) continue;
    const bool pts = g.dim >= 2 && (he == nullptr || he[b].pts != 0u);
    total += static_cast<double>(g.count) * (1.0 + sizeof(T) * (fields + (pts ? g.dim : 0)));
  }
  return total;
}

template <class T>
void enqueue_element_op(hipStream_t s, const Geom& g, int op, const ElemDev* el, const uint8_t* lo, const T* t,
                        const T* o, const int32_t* rung_win, const int32_t* rung_max, int K_max, const T* in, T* out0,
                        T* out1, T* rungs0, T* rungs1, T* pairs, const int32_t* on, GatherExtra<T> extra) {
  if (K_max < 1) return;
  int32_t mw = 0;
  for (int j = 0; j < K_max; ++j) mw = std::max(mw, rung_max[j]);
  const int64_t F = g.batch * g.count;
  const bool timed = timer_family(kTimeGather);
  hipEvent_t t0 = nullptr;
  const auto timed_gather = [&](auto&& launch) {
    if (timed) timer_note(s, &t0);
    launch();
    if (timed) timer_done(s, t0, gather_bytes<T>(g, op, el, on, extra), kTagGather);
  };
  if (op == 3) {
    enqueue_cube_rungs<T>(s, g, CubeOp::kBox, in, rungs0, nullptr, rung_win, K_max, F, mw, nullptr, rungs1);
    timed_gather([&] { launch_gather_all<T, 3>(s, g, el, lo, t, o, rungs0, nullptr, out0, nullptr, on, extra); });
  } else if (op == 2) {
    if (enqueue_cube_rungs<T>(s, g, CubeOp::kMaxMin, in, rungs0, rungs1, rung_win, K_max, F, mw, pairs))
      timed_gather([&] { launch_gather_pairs<T>(s, g, el, lo, t, o, pairs, out0, out1, on, extra); });
    else
      timed_gather([&] { launch_gather_all<T, 2>(s, g, el, lo, t, o, rungs0, rungs1, out0, out1, on, extra); });
  } else {
    enqueue_cube_rungs<T>(s, g, CubeOp::kMax, in, rungs0, nullptr, rung_win, K_max, F, mw, nullptr, rungs1);
    timed_gather([&] { launch_gather_all<T, 0>(s, g, el, lo, t, o, rungs0, nullptr, out0, nullptr, on, extra); });
  }
}

template void enqueue_element_build<float>(hipStream_t, const Geom&, const MethodCtl&, const float*, const float*,
                                           const float*, float*, const int32_t*, uint64_t*);
template void enqueue_element_build<double>(hipStream_t, const Geom&, const MethodCtl&, const double*, const double*,
                                            const double*, double*, const int32_t*, uint64_t*);
template void enqueue_element_weights<float>(hipStream_t, const Geom&, const MethodCtl&, const float*, const float*,
                                             const ElemDev*, uint8_t*, float*, const int32_t*, uint32_t*);
template void enqueue_element_weights<double>(hipStream_t, const Geom&, const MethodCtl&, const double*, const double*,
                                              const ElemDev*, uint8_t*, double*, const int32_t*, uint32_t*);
template void enqueue_element_op<float>(hipStream_t, const Geom&, int, const ElemDev*, const uint8_t*, const float*,
                                        const float*, const int32_t*, const int32_t*, int, const float*, float*, float*,
                                        float*, float*, float*, const int32_t*, GatherExtra<float>);
template void enqueue_element_op<double>(hipStream_t, const Geom&, int, const ElemDev*, const uint8_t*, const double*,
                                         const double*, const int32_t*, const int32_t*, int, const double*, double*,
                                         double*, double*, double*, double*, const int32_t*, GatherExtra<double>);

}  // namespace eda
