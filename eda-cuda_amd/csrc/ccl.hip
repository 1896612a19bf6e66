// SPDX-License-Identifier: Apache-2.0
// Face-connected components (scipy.ndimage.label's default structure; the domain border does not
// connect), by lock-free union-find (Playne and Hawick's scheme; reference
// cpp/src/method/stats_scale.cu:1-83): every voxel of the mask links with its +1 neighbour along
// every axis, a link always hangs the larger root under the smaller one by atomicMin, so every
// parent is smaller than its child and a component's root is its smallest in-element index whatever
// the order of the atomics; a last pass writes every voxel's root.
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

}  // namespace

void enqueue_ccl(hipStream_t s, const Geom& g, const uint8_t* mask, uint8_t bits, int32_t* labels, const int32_t* on) {
  each(s, g, on, [=] __device__(int64_t b, int64_t i) {
    labels[b * g.count + i] = (mask[b * g.count + i] & bits) != 0 ? static_cast<int32_t>(i) : -1;
  });
  each(s, g, on, [=] __device__(int64_t b, int64_t i) {
    int32_t* L = labels + b * g.count;
    const int32_t v = static_cast<int32_t>(i);
    if (L[v] < 0) return;
    for (int a = 0; a < g.dim; ++a) {
      if ((i / g.st[a]) % g.n[a] + 1 >= g.n[a]) continue;
      const int32_t u = v + static_cast<int32_t>(g.st[a]);
      if (L[u] < 0) continue;
      unite(L, v, u);
    }
  });
  each(s, g, on, [=] __device__(int64_t b, int64_t i) {
    int32_t* L = labels + b * g.count;
    const int32_t v = static_cast<int32_t>(i);
    if (L[v] < 0) return;
    const int32_t root = find_root(L, v);
    reinterpret_cast<volatile int32_t*>(L)[v] = root;
  });
}

}  // namespace eda
