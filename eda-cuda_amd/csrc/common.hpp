// SPDX-License-Identifier: Apache-2.0
// Shared pieces of the gfx950 decomposition library: the field layout, canonical scalar arithmetic,
// error handling and the launch helpers every kernel file uses.
//
// Layout. A call's working field is C-contiguous (B, n_0 .. n_{D-1}) with no padding: element b's
// samples start at b * count. The reference pads the last axis to a multiple of 32
// (cpp/src/field/geom.hpp:9-23); the padding changes no result, and wave64 rows of 512 samples are
// aligned either way, so the library keeps the dense layout and never tests for pad samples.
//
// Arithmetic. Every floating-point operation that feeds a result is a single IEEE operation in a
// fixed order. All HIP code is compiled with -ffp-contract=off, so the compiler never fuses a
// multiply and an add; the order of operations is the one the reference's canonical intrinsics fix
// (cpp/src/field/common.cuh:1-131), so the float32 and float64 kernels reproduce the CPU oracle's
// bits wherever the oracle's arithmetic is the same (everything but libm and the DCT sums).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>

namespace eda {

inline constexpr int kMaxDim = 3;
inline constexpr int kMaxRungs = 32;
inline constexpr int kThreads = 256;

// ------------------------------------------------------------------------------------------------
// Errors: every HIP call is checked; a failure becomes eda::HipError (the C ABI maps it to a return
// code, Python to eda.CudaError / MemoryError).
struct HipError : std::runtime_error {
  hipError_t code;
  HipError(hipError_t c, const std::string& what) : std::runtime_error(what), code(c) {}
};
struct OutOfDeviceMemory : std::runtime_error {
  using std::runtime_error::runtime_error;
};

[[noreturn]] void throw_hip(hipError_t e, const char* expr, const char* file, int line);
#define EDA_HIP(expr)                                                 \
  do {                                                                \
    const hipError_t eda_e_ = (expr);                                 \
    if (eda_e_ != hipSuccess) ::eda::throw_hip(eda_e_, #expr, __FILE__, __LINE__); \
  } while (0)
// Issues the parameter uploads the calling thread has staged since the last call (as one copy, on
// the stream they were staged for). Every kernel launch and every device-side copy that may read
// an upload runs it first (EDA_GO wraps the launches).
void flush_uploads();
#define EDA_GO(...)                 \
  do {                              \
    ::eda::flush_uploads();         \
    hipLaunchKernelGGL(__VA_ARGS__); \
  } while (0)

#define EDA_LAUNCH(name)                                              \
  do {                                                                \
    const hipError_t eda_e_ = hipGetLastError();                      \
    if (eda_e_ != hipSuccess) ::eda::throw_hip(eda_e_, name, __FILE__, __LINE__); \
  } while (0)

// ------------------------------------------------------------------------------------------------
// Geometry of a batch of fields.
struct Geom {
  int32_t dim;           // D in 1..3
  int32_t n[kMaxDim];    // extents (1 beyond dim)
  int64_t batch;         // B
  int64_t count;         // samples per element
  int64_t st[kMaxDim];   // element strides (st[D-1] = 1)
};
inline Geom make_geom(int dim, int64_t batch, const int64_t* n) {
  Geom g{};
  g.dim = dim;
  g.batch = batch;
  g.count = 1;
  for (int a = 0; a < kMaxDim; ++a) g.n[a] = 1;
  for (int a = 0; a < dim; ++a) {
    g.n[a] = static_cast<int32_t>(n[a]);
    g.count *= n[a];
  }
  int64_t acc = 1;
  for (int a = kMaxDim - 1; a >= 0; --a) {
    if (a >= dim) {
      g.st[a] = 1;
      continue;
    }
    g.st[a] = acc;
    acc *= g.n[a];
  }
  return g;
}
inline int64_t field_elems(const Geom& g) { return g.batch * g.count; }

// ------------------------------------------------------------------------------------------------
// Canonical scalar helpers (cpp/src/field/common.cuh:62-131, method.hpp:172-178).
template <class T>
__host__ __device__ __forceinline__ T max_c(T a, T b) {
  return (a > b || (a == b && !signbit(a))) ? a : b;
}
template <class T>
__host__ __device__ __forceinline__ T min_c(T a, T b) {
  return (a < b || (a == b && signbit(a))) ? a : b;
}
__host__ __device__ __forceinline__ int64_t floor_div(int64_t a, int64_t b) {
  const int64_t q = a / b;
  return (a % b != 0 && a < 0) ? q - 1 : q;
}
__host__ __device__ __forceinline__ int64_t floor_mod(int64_t a, int64_t b) {
  const int64_t m = a % b;
  return m < 0 ? m + b : m;
}
__host__ __device__ __forceinline__ int32_t reflect_index(int32_t e, int32_t n) {
  return e < 0 ? -e - 1 : (e >= n ? 2 * n - 1 - e : e);
}
__host__ __device__ __forceinline__ int64_t reflect_any(int64_t i, int64_t n) {
  const int64_t m = floor_mod(i, 2 * n);
  return m < n ? m : 2 * n - 1 - m;
}
__host__ __device__ __forceinline__ int32_t odd_window(double v) {
  const double u = v > 1.0 ? v : 1.0;
  return 2 * static_cast<int32_t>(floor(u / 2.0)) + 1;
}
template <class T>
__host__ __device__ __forceinline__ T clip_np(T v, T lo, T hi) {
  const T a = v >= lo ? v : lo;
  return a <= hi ? a : hi;
}
__host__ __device__ __forceinline__ uint64_t dbits(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
  return static_cast<uint64_t>(__double_as_longlong(d));
#else
  uint64_t u;
  std::memcpy(&u, &d, 8);
  return u;
#endif
}
__host__ __device__ __forceinline__ double bits_d(uint64_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __longlong_as_double(static_cast<long long>(u));
#else
  double d;
  std::memcpy(&d, &u, 8);
  return d;
#endif
}
template <class T>
__host__ __device__ __forceinline__ uint64_t abs_bits(T v) {
  return dbits(static_cast<double>(v)) & 0x7fffffffffffffffULL;
}
template <class T>
__device__ __forceinline__ T inf_t() {
  return __builtin_huge_val();
}
template <>
__device__ __forceinline__ float inf_t<float>() {
  return __builtin_huge_valf();
}

// Integer order keys of floating values: the integer order of key(v) is the order of max_c/min_c
// on non-NaN values (-0 below +0); the map is its own inverse.
__device__ __forceinline__ int32_t okey(float v) {
  const int32_t b = __float_as_int(v);
  return b ^ ((b >> 31) & 0x7fffffff);
}
__device__ __forceinline__ float okey_inv(int32_t k) { return __int_as_float(k ^ ((k >> 31) & 0x7fffffff)); }
__device__ __forceinline__ long long okey(double v) {
  const long long b = __double_as_longlong(v);
  return b ^ ((b >> 63) & 0x7fffffffffffffffLL);
}
__device__ __forceinline__ double okey_inv(long long k) {
  return __longlong_as_double(k ^ ((k >> 63) & 0x7fffffffffffffffLL));
}
template <class T>
using okey_t = std::conditional_t<sizeof(T) == 4, int32_t, long long>;

// ------------------------------------------------------------------------------------------------
// Launch helpers. each(): a grid over (blocks of samples, elements); f(b, i) for element b, sample
// i < count of every element whose run flag on[b] is non-zero (on nullptr: all).
template <class F>
__global__ void __launch_bounds__(kThreads) k_each(int64_t count, int64_t batch, const int32_t* __restrict__ on, F f) {
  for (int64_t b = blockIdx.y; b < batch; b += gridDim.y) {
    if (on != nullptr && on[b] == 0) continue;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(kThreads) + threadIdx.x; i < count;
         i += static_cast<int64_t>(gridDim.x) * kThreads)
      f(b, i);
  }
}
template <class F>
void each(hipStream_t s, const Geom& g, const int32_t* on, F f) {
  int64_t gx = (g.count + kThreads - 1) / kThreads;
  gx = gx < 1 ? 1 : (gx > 8192 ? 8192 : gx);
  const int64_t gy = g.batch < 65535 ? g.batch : 65535;
  EDA_GO(k_each<F>, dim3(static_cast<unsigned>(gx), static_cast<unsigned>(gy)), dim3(kThreads), 0, s,
                     g.count, g.batch, on, f);
  EDA_LAUNCH("k_each");
}
// A flat grid over n items: f(i).
template <class F>
__global__ void __launch_bounds__(kThreads) k_flat(int64_t n, F f) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kThreads) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kThreads)
    f(i);
}
template <class F>
void flat(hipStream_t s, int64_t n, F f) {
  if (n <= 0) return;
  int64_t gx = (n + kThreads - 1) / kThreads;
  gx = gx > 65536 ? 65536 : gx;
  EDA_GO(k_flat<F>, dim3(static_cast<unsigned>(gx)), dim3(kThreads), 0, s, n, f);
  EDA_LAUNCH("k_flat");
}

// Wave64 reductions.
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t u = __shfl_xor(v, o, 64);
    v = u > v ? u : v;
  }
  return v;
}
__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t u = __shfl_xor(v, o, 64);
    v = u < v ? u : v;
  }
  return v;
}
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ uint32_t wave_or_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}

// Workgroup i runs on XCD i mod 8. The item of workgroup i when chunks of c consecutive items
// rotate over the XCDs (the grid a multiple of 8 c): the tiles of one chunk are neighbours in
// memory and share an XCD's L2, and the items of any few elements spread over all eight XCDs (a
// launch in which only some elements run keeps the whole chip busy).
__device__ __forceinline__ int64_t xcd_item(int64_t i, int32_t c) {
  const int64_t j = i >> 3;
  return ((j / c) * 8 + (i & 7)) * c + j % c;
}
// The largest chunk of 8, 4, 2 items whose rotation tiles a grid of wgs workgroups (0: none).
inline int32_t xcd_chunk(int64_t wgs) {
  for (int32_t c = 8; c >= 2; c >>= 1)
    if (wgs % (8 * c) == 0) return c;
  return 0;
}

}  // namespace eda
