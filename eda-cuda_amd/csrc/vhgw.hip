// SPDX-License-Identifier: Apache-2.0
// The separable cube filter: running max, min or box mean of an odd window along every axis, by
// van Herk / Gil-Werman passes (reference: cpp/src/field/vhgw.cu:1-36 for the canonical order,
// cpp/src/method/common.cpp:65-136 for the cube; host twin cpp/tests/twin.hpp:74-133).
//
// Canonical arithmetic of one line of n samples and window w: r = min((w-1)/2, n-1), wa = 2r+1;
// the extended line e in [-r, n+r) reads the line through half-sample reflection; it is cut into
// segments of wa samples starting at o + k wa, o = (n (wa+1)/2) mod wa. The prefix fold G of a
// segment runs in 32-sample chunks: a sequential fold inside a chunk and a sequential carry over the
// chunk totals, G = carry op local (chunk 0: local). The suffix fold H is G of the reversed line.
// y(x) = op(H(x-r), G(x+r)); an aligned window (both ends in one segment) of a sum takes
// 0.5 (H + G). Max and min are exact, so they fold integer order keys; sums depend only on (line,
// n, w), never on the launch shape.
//
// gfx950 mapping. A workgroup of 256 threads (4 waves of 64) stages P lines of one element in LDS
// (V: the input line, R: the H folds and then the output), P = 4096 / n lines for the contiguous
// axis and P adjacent columns for the strided axes, so that every global load and store is a
// coalesced run of the row. The folds run as chunk tasks (line, segment, chunk): phase A folds
// each chunk (32 sequential LDS operations) and records its total, phase B turns the totals of a
// segment into carries by an in-place exclusive scan, phase C applies them. Line storage uses the
// XOR swizzle i ^ ((i >> 5) & 31) (float) / & 15 (8-byte) and a line stride of round_up(n, 32) + 1,
// so that chunk walks (32 lanes 32 samples apart) and column loads hit distinct LDS banks.
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace eda {
namespace {

constexpr int kVMax = 0, kVMin = 1, kVSum = 2;

template <class S>
__device__ __forceinline__ int32_t swz(int32_t i) {
  constexpr int32_t m = sizeof(S) == 4 ? 31 : 15;
  return i ^ ((i >> 5) & m);
}
__host__ __device__ __forceinline__ int32_t line_stride(int32_t n) { return (n + 31) / 32 * 32 + 1; }

template <int OP, class S>
__device__ __forceinline__ S fold_op(S a, S b) {
  if constexpr (OP == kVMax) return a > b ? a : b;
  else if constexpr (OP == kVMin) return a < b ? a : b;
  else return a + b;
}

// The per-window constants of a line (all lines of a workgroup share the element and its window).
struct WinInfo {
  int32_t r, wa, o, k_lo, nsegs, cs, tasks;
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
  v.tasks = v.nsegs * v.cs;
  return v;
}

struct VPass {
  Geom g;
  int32_t axis;
  int32_t n;            // line length
  int64_t step;         // sample stride along the line
  int64_t inner;        // adjacent lines (1 on the contiguous axis)
  int64_t outer;        // line groups per element along the other axes
  int32_t P;            // lines per workgroup
  int64_t groups;       // workgroups per element
  int32_t tmax;         // chunk-task slots per line (LDS)
  int32_t nf;           // fields: 1, or 2 (max then min, same input, or two inputs)
  int32_t same_input;   // nf == 2 from one input (staged once)
  int32_t ops[2];
  int32_t scale;        // multiply by fl_T(1 / prod of the clamped sides) (the last box pass)
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

// One direction of one stage on every line of the workgroup. DIR 1: H (the reversed line), whose
// local folds and then carried folds go into R at the output slots; DIR 0: G, which combines with
// R's H into the stage output in R.
template <class S, int OP, int DIR>
__device__ void stage_dir(const S* __restrict__ V, S* __restrict__ R, S* __restrict__ tot, int P, int32_t ls,
                          int32_t tmax, int32_t n, const WinInfo& wi) {
  const int tid = threadIdx.x;
  const int32_t r = wi.r, wa = wi.wa;
  const int32_t total = P * wi.tasks;
  const bool chain = wi.cs > 1;
  const auto load = [&](const S* Vl, int32_t e) {
    int32_t i = reflect_index(e, n);
    if (DIR) i = n - 1 - i;
    return Vl[swz<S>(i)];
  };
  const auto slot = [&](int32_t e) {
    int32_t x = e - r;
    if (DIR) x = n - 1 - x;
    return swz<S>(x);
  };
  // phase A: chunk folds (H: local values into R at the output slots) and the chunk totals
  if (DIR == 1 || chain) {
    for (int32_t t = tid; t < total; t += kThreads) {
      const int32_t p = t / wi.tasks, tl = t - p * wi.tasks;
      const int32_t k = wi.k_lo + tl / wi.cs, q = tl % wi.cs;
      const int32_t s = wi.o + k * wa;
      const int32_t e_last = min(s + wa - 1, n - 1 + r);
      const int32_t c0 = s + 32 * q;
      const int32_t c1 = min(c0 + 31, e_last);
      if (c0 > c1) continue;
      const S* Vl = V + p * ls;
      S* Rl = R + p * ls;
      S acc = load(Vl, c0);
      if (DIR == 1 && c0 >= r) Rl[slot(c0)] = acc;
      for (int32_t e = c0 + 1; e <= c1; ++e) {
        acc = fold_op<OP>(acc, load(Vl, e));
        if (DIR == 1 && e >= r) Rl[slot(e)] = acc;
      }
      if (chain) tot[p * tmax + tl] = acc;
    }
    __syncthreads();
  }
  // phase B: per segment, the exclusive scan of the chunk totals (in place: slot q holds the carry
  // of chunk q, the fold of the totals of chunks 0 .. q-1)
  if (chain) {
    for (int32_t t = tid; t < P * wi.nsegs; t += kThreads) {
      const int32_t p = t / wi.nsegs, kk = t - p * wi.nsegs;
      const int32_t s = wi.o + (wi.k_lo + kk) * wa;
      const int32_t e_last = min(s + wa - 1, n - 1 + r);
      S* tl = tot + p * tmax + kk * wi.cs;
      S run = tl[0];
      for (int32_t q = 1; q < wi.cs && s + 32 * q <= e_last; ++q) {
        const S v = tl[q];
        tl[q] = run;
        run = fold_op<OP>(run, v);
      }
    }
    __syncthreads();
  }
  // phase C
  for (int32_t t = tid; t < total; t += kThreads) {
    const int32_t p = t / wi.tasks, tl = t - p * wi.tasks;
    const int32_t k = wi.k_lo + tl / wi.cs, q = tl % wi.cs;
    const int32_t s = wi.o + k * wa;
    const int32_t e_last = min(s + wa - 1, n - 1 + r);
    const int32_t c0 = s + 32 * q;
    const int32_t c1 = min(c0 + 31, e_last);
    if (c0 > c1) continue;
    S* Rl = R + p * ls;
    const S carry = q > 0 ? tot[p * tmax + tl] : S(0);
    if constexpr (DIR == 1) {
      if (q == 0) continue;
      for (int32_t e = max(c0, r); e <= c1; ++e) {
        const int32_t sl = slot(e);
        Rl[sl] = fold_op<OP>(carry, Rl[sl]);
      }
    } else {
      const S* Vl = V + p * ls;
      const int32_t e_al = s + wa - 1;
      S loc = S(0);
      for (int32_t e = c0; e <= c1; ++e) {
        const S xv = load(Vl, e);
        loc = e == c0 ? xv : fold_op<OP>(loc, xv);
        if (e >= r) {
          const S gv = q > 0 ? fold_op<OP>(carry, loc) : loc;
          const int32_t sl = slot(e);
          const S h = Rl[sl];
          if constexpr (OP == kVSum) {
            const S sum = h + gv;
            Rl[sl] = e == e_al ? sum * S(0.5) : sum;
          } else {
            Rl[sl] = fold_op<OP>(h, gv);
          }
        }
      }
    }
  }
  __syncthreads();
}

template <class S, int OP>
__device__ void run_stage(const S* V, S* R, S* tot, int P, int32_t ls, int32_t tmax, int32_t n, const WinInfo& wi) {
  stage_dir<S, OP, 1>(V, R, tot, P, ls, tmax, n, wi);
  stage_dir<S, OP, 0>(V, R, tot, P, ls, tmax, n, wi);
}

// S: the LDS type (T for sums; the integer order key of T for max / min).
template <class T, class S>
__global__ void __launch_bounds__(kThreads) k_vhgw(VPass a, VPtrs<T> ptr) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int64_t wg = blockIdx.x;
  const int64_t b = wg / a.groups;
  const int64_t grp = wg - b * a.groups;
  const int64_t rung = blockIdx.y;
  const int32_t w = ptr.win[rung * a.g.batch + b];
  if (w < 1) return;  // uniform: the element does not run
  const int32_t n = a.n;
  const int32_t ls = line_stride(n);
  const WinInfo wi = win_info(n, w);
  S* V = reinterpret_cast<S*>(smem);
  S* R = V + a.P * ls;
  S* tot = R + a.P * ls;
  // the lines of this workgroup: P consecutive rows (contiguous axis) or P adjacent columns
  int64_t base0;
  int32_t lines;
  const bool contiguous = a.inner == 1;
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
  const int32_t total_pos = lines * n;
  T fac = T(1);
  if (a.scale) {
    double prod = 1.0;
    for (int ax = 0; ax < a.g.dim; ++ax) {
      const int32_t h = (w - 1) / 2;
      prod *= static_cast<double>(2 * min(h, a.g.n[ax] - 1) + 1);
    }
    fac = static_cast<T>(1.0 / prod);
  }
  for (int f = 0; f < a.nf; ++f) {
    const int op = a.ops[f];
    if (f == 0 || !a.same_input) {
      const T* src = ptr.in[f] + rung * ptr.in_stride;
      for (int32_t idx = threadIdx.x; idx < total_pos; idx += kThreads) {
        int32_t p, i;
        int64_t at;
        if (contiguous) {
          p = idx / n;
          i = idx - p * n;
          at = base0 + static_cast<int64_t>(p) * n + i;
        } else {
          i = idx / lines;
          p = idx - i * lines;
          at = base0 + static_cast<int64_t>(i) * a.step + p;
        }
        const T v = src[at];
        S sv;
        if constexpr (std::is_same_v<S, T>) sv = v;
        else sv = okey(v);
        V[p * ls + swz<S>(i)] = sv;
      }
      __syncthreads();
    }
    if constexpr (std::is_same_v<S, T>) {
      run_stage<S, kVSum>(V, R, tot, lines, ls, a.tmax, n, wi);
    } else {
      if (op == kVMax) run_stage<S, kVMax>(V, R, tot, lines, ls, a.tmax, n, wi);
      else run_stage<S, kVMin>(V, R, tot, lines, ls, a.tmax, n, wi);
    }
    T* dst = ptr.out[f] + rung * ptr.out_stride;
    for (int32_t idx = threadIdx.x; idx < total_pos; idx += kThreads) {
      int32_t p, i;
      int64_t at;
      if (contiguous) {
        p = idx / n;
        i = idx - p * n;
        at = base0 + static_cast<int64_t>(p) * n + i;
      } else {
        i = idx / lines;
        p = idx - i * lines;
        at = base0 + static_cast<int64_t>(i) * a.step + p;
      }
      const S sv = R[p * ls + swz<S>(i)];
      T y;
      if constexpr (std::is_same_v<S, T>) y = sv;
      else y = okey_inv(sv);
      if (a.scale) y = y * fac;
      dst[at] = y;
    }
    __syncthreads();
  }
}

template <class T>
size_t lds_bytes(int32_t n, int P, int32_t tmax) {
  const size_t s = sizeof(T);
  return (2 * static_cast<size_t>(P) * line_stride(n) + static_cast<size_t>(P) * tmax) * s;
}

}  // namespace

// The chunk-task slots per line for the windows of a launch.
int32_t vhgw_task_slots(int32_t n, int32_t max_w) {
  // tasks(w) is not monotone in w: take the largest over every odd w up to max_w, by its closed form
  int32_t best = 1;
  const int32_t wmax = std::min<int32_t>(max_w, 2 * n - 1);
  for (int32_t w = 33; w <= wmax; w += 2) best = std::max(best, win_info(n, w).tasks);
  return best;
}

template <class T>
void enqueue_vhgw(hipStream_t s, const Geom& g, int axis, int nf, bool same_input, const int* ops, bool scale,
                  const T* const* in, T* const* out, const int32_t* win, int32_t max_w, int32_t rungs,
                  int64_t in_stride, int64_t out_stride) {
  if (max_w < 1 || rungs < 1) return;
  if (rungs > 65535) throw std::invalid_argument("eda: too many rungs for one launch");
  VPass a{};
  a.g = g;
  a.axis = axis;
  a.n = g.n[axis];
  a.step = g.st[axis];
  a.inner = axis == g.dim - 1 ? 1 : g.st[axis];
  a.outer = g.count / (static_cast<int64_t>(a.n) * a.inner);  // rows (contiguous axis) or column groups
  const int32_t n = a.n;
  int P;
  if (a.inner == 1) {
    P = std::max(1, std::min(16, 4096 / n));
    a.groups = (a.outer + P - 1) / P;
  } else {
    P = static_cast<int>(std::min<int64_t>(a.inner, std::max(1, std::min(8, 8192 / n))));
    const int64_t gpo = (a.inner + P - 1) / P;
    a.groups = a.outer * gpo;
  }
  a.P = P;
  a.tmax = vhgw_task_slots(n, max_w);
  a.nf = nf;
  a.same_input = same_input ? 1 : 0;
  a.ops[0] = ops[0];
  a.ops[1] = nf > 1 ? ops[1] : ops[0];
  a.scale = scale ? 1 : 0;
  VPtrs<T> p{};
  p.in[0] = in[0];
  p.in[1] = nf > 1 ? in[1] : in[0];
  p.out[0] = out[0];
  p.out[1] = nf > 1 ? out[1] : out[0];
  p.win = win;
  p.in_stride = in_stride;
  p.out_stride = out_stride;
  const int64_t wgs = g.batch * a.groups;
  const bool sum = ops[0] == kVSum;
  const size_t lds = lds_bytes<T>(n, P, a.tmax);
  hipEvent_t t0 = nullptr;
  const bool timed = timer_on();
  if (timed) timer_note(s, &t0);
  const dim3 grid(static_cast<unsigned>(wgs), static_cast<unsigned>(rungs));
  if (sum) {
    auto k = k_vhgw<T, T>;
    if (lds > 65536) EDA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    hipLaunchKernelGGL(k, grid, dim3(kThreads), lds, s, a, p);
  } else {
    auto k = k_vhgw<T, okey_t<T>>;
    if (lds > 65536) EDA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    hipLaunchKernelGGL(k, grid, dim3(kThreads), lds, s, a, p);
  }
  EDA_LAUNCH("k_vhgw");
  if (timed) {
    int64_t active = mirror_active(win, g.batch * rungs);
    if (active < 0) active = g.batch * rungs;
    const int fields = (nf == 2 && !same_input ? 2 : 1) + nf;
    timer_done(s, t0, static_cast<double>(fields) * active * g.count * sizeof(T));
  }
}

// The cube (reference cpp/src/method/common.cpp:65-136): box runs axis 0 first, the last pass
// scaled; max / min run the last axis first; maxmin stages the input once for both outputs. With
// rungs > 1 every pass runs all rungs in one launch: rung j filters in with the windows win + j B
// into out0 + j stride (out1 + j stride).
template <class T>
void enqueue_cube_rungs(hipStream_t s, const Geom& g, CubeOp op, const T* in, T* out0, T* out1, const int32_t* win,
                        int32_t rungs, int64_t stride, int32_t max_w) {
  if (max_w < 1) return;
  const int D = g.dim;
  if (op == CubeOp::kBox) {
    for (int a = 0; a < D; ++a) {
      const int ops[1] = {kVSum};
      const T* src[1] = {a == 0 ? in : out0};
      T* dst[1] = {out0};
      enqueue_vhgw<T>(s, g, a, 1, true, ops, a == D - 1, src, dst, win, max_w, rungs, a == 0 ? 0 : stride, stride);
    }
    return;
  }
  for (int a = D - 1; a >= 0; --a) {
    const bool first = a == D - 1;
    const int64_t ins = first ? 0 : stride;
    if (op == CubeOp::kMaxMin) {
      const int ops[2] = {kVMax, kVMin};
      const T* src[2] = {first ? in : out0, first ? in : out1};
      T* dst[2] = {out0, out1};
      enqueue_vhgw<T>(s, g, a, 2, first, ops, false, src, dst, win, max_w, rungs, ins, stride);
    } else {
      const int ops[1] = {op == CubeOp::kMax ? kVMax : kVMin};
      const T* src[1] = {first ? in : out0};
      T* dst[1] = {out0};
      enqueue_vhgw<T>(s, g, a, 1, true, ops, false, src, dst, win, max_w, rungs, ins, stride);
    }
  }
}

template <class T>
void enqueue_cube(hipStream_t s, const Geom& g, CubeOp op, const T* in, T* out0, T* out1, const int32_t* win,
                  int32_t max_w) {
  enqueue_cube_rungs<T>(s, g, op, in, out0, out1, win, 1, 0, max_w);
}

template void enqueue_cube<float>(hipStream_t, const Geom&, CubeOp, const float*, float*, float*, const int32_t*, int32_t);
template void enqueue_cube<double>(hipStream_t, const Geom&, CubeOp, const double*, double*, double*, const int32_t*,
                                   int32_t);
template void enqueue_cube_rungs<float>(hipStream_t, const Geom&, CubeOp, const float*, float*, float*, const int32_t*,
                                        int32_t, int64_t, int32_t);
template void enqueue_cube_rungs<double>(hipStream_t, const Geom&, CubeOp, const double*, double*, double*,
                                         const int32_t*, int32_t, int64_t, int32_t);

}  // namespace eda
