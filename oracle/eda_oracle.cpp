// SPDX-License-Identifier: Apache-2.0
//
// THE ORACLE. Test infrastructure only: a CPU restatement of the decomposition method of
// giorgioangel/eda-cuda 0.1.0 (`eda.decompose`), written from the reference's source text, used by
// tests/, by __graft_entry__.smoke() and by bench.py's CPU-baseline leg as the checker. It is never
// linked into, loaded by or called from the product library (eda-cuda_amd/), which fails loudly
// without its HIP build instead of falling back to this code.
//
// Templated on the arithmetic type T (float or double). T = double is "the FP64 oracle"; T = float
// mirrors the float32 method so that the GPU's float32 primitives can be compared bit for bit.
// Every floating-point operation that feeds a result is one IEEE operation in a fixed order; the
// file is compiled with -ffp-contract=off and without fast-math, so the compiler neither fuses nor
// reorders them.
//
// Elements of a batch are independent (the reference's results do not depend on the batch, the
// chunking or the launch shape: cpp/src/method/method.hpp:36-39), so the oracle runs every element
// through its own mode loop; OpenMP splits the batch over threads without changing any bit.
//
// Parity pins (what checks each function; tests/ holds the checks):
//   vhgw_line / cube          brute-force sliding max/min/mean with half-sample reflection
//                              (scipy.ndimage.*_filter1d mode='reflect'), exact for max/min,
//                              within rounding for the box; golden corpus end to end.
//   row_energy / amax         closed form (numpy sums) within rounding; amax exact.
//   select_median/quantile    numpy.median / numpy.quantile(method='linear'), and the NumPy
//                              quantile table of cpp/tests/test_stats.cu:282-324 (bitwise).
//   noise_sigma               closed form 1.4826 MAD / sqrt 2 with numpy.
//   sig_flags                 scipy.signal.find_peaks + peak_prominences.
//   ccl                       scipy.ndimage.label after canonical relabelling.
//   dct / riesz               scipy.fft.dctn / idctn(norm='ortho') and DCT-basis plane waves.
//   eigen                     numpy.linalg.eigh.
//   monogenic                 DCT-basis fields (exact amplitude 1, phase, frequency, direction,
//                              coherence 1) and a NumPy FFT computation of the same definitions.
//   rung_ladder               the KAT cpp/tests/rungs_table.inc (66 rows).
//   interp (gather)           scipy.ndimage.map_coordinates(order=1) on the reflected extension.
//   decompose (T = double)    the golden corpus tests/data/golden (9 FP64 cases: decisions exact,
//                              rows within the corpus's own bound), and the decisions of the
//                              parity manifest tests/data/parity_v2.json (143 float64 cases).
//   decompose (T = float)     the decisions of the parity manifest (142 of its 144 float32 cases
//                              on the CPU; the other two run on the GPU).
// The G2/G3 stages, the scale field, the floor and the sift kernels are pinned only end to end
// (golden corpus: g3blobs_2d_64 exercises G2/G3) and by invariants (reconstruction, gain 2^j).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <set>
#include <stdexcept>
#include <type_traits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

constexpr int kMaxRungs = 32;       // cpp/src/method/method.hpp:61
constexpr int kMaxG2Rungs = 32;     // method.hpp:64
constexpr double kRungRatio = 1.4142135623730951;  // method.hpp:66

// ------------------------------------------------------------------------------------------------
// Per-call constants: cpp/src/method/common.cpp:14-63.
struct Ctl {
  int32_t n_max;
  int64_t count;
  int32_t sifts, u, s, sig_passes, growth_from, anorm_smooth;
  double kappa, beta, c_lo, c_hi, reset, q, span, growth;
  double sig_theta, sig_rho, sig_noise;
  double absent, absent_jump, absent_reach, absent_lo;
  double t0, t1, anorm_floor, ramp_k0, ramp_k1;
  double fine_beta, fine_theta, fine_jump, fine_lo, fine_pop, fine_core, split;
  double tiny;
};

Ctl make_ctl(int D, const int32_t* n, int64_t count, bool f32) {
  Ctl c{};
  c.n_max = 1;
  for (int a = 0; a < D; ++a) c.n_max = std::max(c.n_max, n[a]);
  c.count = count;
  c.sifts = 3;
  c.kappa = 1.0;
  c.beta = 2.0;
  c.c_lo = 1.5;
  c.c_hi = 2.0;
  c.reset = 2.0;
  c.q = 4.0;
  c.span = 1.0;
  c.growth = 2.0;
  c.growth_from = 2;
  c.sig_theta = 0.3;
  c.sig_rho = 2.0;
  c.sig_noise = 1.0;
  c.sig_passes = 2;
  c.absent = 3.0;
  c.absent_jump = 2.0;
  c.absent_reach = 1.0;
  c.absent_lo = 1.2;
  c.t0 = 0.2;
  c.t1 = 0.5;
  c.u = 2;
  c.s = 1;
  c.anorm_floor = 0.1;
  c.anorm_smooth = 1;
  c.ramp_k0 = 0.3;
  c.ramp_k1 = 0.6;
  c.fine_beta = 2.0;
  c.fine_theta = 1.0;
  c.fine_jump = 2.0;
  c.fine_lo = 1.5;
  c.fine_pop = 0.25;
  c.fine_core = 2.0;
  c.split = 1.0;
  c.tiny = f32 ? static_cast<double>(FLT_MIN) : 1e-300;
  return c;
}

// ------------------------------------------------------------------------------------------------
// Scalar helpers: cpp/src/field/common.cuh:24-131, method.hpp:172-178.

template <class T>
inline T max_c(T a, T b) { return (a > b || (a == b && !std::signbit(a))) ? a : b; }
template <class T>
inline T min_c(T a, T b) { return (a < b || (a == b && std::signbit(a))) ? a : b; }
inline int64_t floor_div(int64_t a, int64_t b) {
  const int64_t q = a / b;
  return (a % b != 0 && a < 0) ? q - 1 : q;
}
inline int64_t floor_mod(int64_t a, int64_t b) {
  const int64_t m = a % b;
  return m < 0 ? m + b : m;
}
inline int32_t reflect_index(int32_t e, int32_t n) { return e < 0 ? -e - 1 : (e >= n ? 2 * n - 1 - e : e); }
inline int64_t reflect_any(int64_t i, int64_t n) {  // element.cuh:72-75
  const int64_t m = floor_mod(i, 2 * n);
  return m < n ? m : 2 * n - 1 - m;
}
inline int32_t odd_window(double v) {
  const double u = v > 1.0 ? v : 1.0;
  return 2 * static_cast<int32_t>(std::floor(u / 2.0)) + 1;
}
inline uint64_t dbits(double d) {
  uint64_t u;
  std::memcpy(&u, &d, 8);
  return u;
}
inline double bits_d(uint64_t u) {
  double d;
  std::memcpy(&d, &u, 8);
  return d;
}
template <class T>
inline uint64_t abs_bits(T v) { return dbits(static_cast<double>(v)) & 0x7fffffffffffffffULL; }
template <class T>
inline double tie_tolerance(double tau, double amax) {  // common.cuh:128-131
  return static_cast<double>(static_cast<T>(tau) * static_cast<T>(amax));
}
template <class T>
inline T clip(T v, T lo, T hi) {  // sift.cu:31-33 (np.max / np.min keep the first operand on ties)
  const T a = v >= lo ? v : lo;
  return a <= hi ? a : hi;
}

// ------------------------------------------------------------------------------------------------
// Geometry of one element: C-order extents, no padding (the reference pads the last axis to 32;
// the padding changes no result: cpp/src/field/geom.hpp:11-13).
struct Shape {
  int D = 1;
  int32_t n[3] = {1, 1, 1};
  int64_t count = 1;
  int64_t st[3] = {1, 1, 1};  // element strides
};
inline Shape make_shape(int D, const int64_t* n) {
  Shape s;
  s.D = D;
  s.count = 1;
  for (int a = 0; a < D; ++a) {
    s.n[a] = static_cast<int32_t>(n[a]);
    s.count *= n[a];
  }
  int64_t acc = 1;
  for (int a = D - 1; a >= 0; --a) {
    s.st[a] = acc;
    acc *= s.n[a];
  }
  return s;
}
inline void coords_of(const Shape& s, int64_t i, int32_t* x) {
  for (int a = s.D - 1; a >= 0; --a) {
    x[a] = static_cast<int32_t>(i % s.n[a]);
    i /= s.n[a];
  }
}
// Visits every line along axis a: f(base) with samples at base + i * st[a], i < n[a].
template <class F>
inline void for_lines(const Shape& s, int a, F&& f) {
  const int64_t n = s.n[a], st = s.st[a];
  const int64_t outer = s.count / (n * st);
  for (int64_t o = 0; o < outer; ++o)
    for (int64_t j = 0; j < st; ++j) f(o * n * st + j);
}

// ------------------------------------------------------------------------------------------------
// The van Herk / Gil-Werman pass (cpp/src/field/vhgw.cu:1-36; host twin cpp/tests/twin.hpp:74-133):
// along one line of n samples and an odd window w, r = min((w-1)/2, n-1), wa = 2r+1; the extended
// line e in [-r, n+r) reads the line by half-sample reflection; segments of wa samples start at
// o + k wa, o = (n (wa+1)/2) mod wa; the prefix fold G of a segment runs in 32-sample chunks (a
// sequential fold inside a chunk, a sequential carry over the chunk totals); the suffix fold H is G
// of the reversed line; y(x) = op(H(x-r), G(x+r)), the sum of an aligned window taking 0.5 (H + G).
enum Op { kMax = 0, kMin = 1, kSum = 2 };

template <class T>
inline T apply(Op op, T a, T b) { return op == kMax ? max_c(a, b) : (op == kMin ? min_c(a, b) : a + b); }

// G at the positions e = r .. n-1+r of the line (rev: of the reversed line) into out[e - r].
template <class T>
void prefix_folds(const T* line, int32_t n, int32_t r, int32_t wa, int32_t o, Op op, bool rev, T* out) {
  const int64_t k0 = floor_div(static_cast<int64_t>(r) - o, wa);
  int32_t s = static_cast<int32_t>(o + k0 * wa);  // start of the segment holding e = r (>= -r)
  T local = T(0), carry = T(0);
  for (int32_t e = s; e <= n - 1 + r; ++e) {
    const int32_t d = e - s;
    int32_t i = reflect_index(e, n);
    if (rev) i = n - 1 - i;
    const T v = line[i];
    const int32_t q = d / 32;
    local = (d % 32 == 0) ? v : apply(op, local, v);
    const T g = q == 0 ? local : apply(op, carry, local);
    if (e >= r) out[e - r] = g;
    if (d % 32 == 31) carry = q == 0 ? local : apply(op, carry, local);
    if (d == wa - 1) s += wa;
  }
}

template <class T>
void vhgw_line(const T* x, int64_t sx, int32_t n, int32_t w, Op op, T* y, int64_t sy, std::vector<T>& buf) {
  const int32_t r = std::min((w - 1) / 2, n - 1);
  const int32_t wa = 2 * r + 1;
  const int32_t o = static_cast<int32_t>((static_cast<int64_t>(n) * ((wa + 1) / 2)) % wa);
  buf.resize(3 * static_cast<size_t>(n));
  T* line = buf.data();
  T* G = line + n;
  T* Hr = G + n;
  for (int32_t i = 0; i < n; ++i) line[i] = x[i * sx];
  prefix_folds(line, n, r, wa, o, op, false, G);
  prefix_folds(line, n, r, wa, o, op, true, Hr);
  for (int32_t i = 0; i < n; ++i) {
    const T h = Hr[n - 1 - i];  // H(i - r) = G_rev(n - 1 - (i - r))
    const T g = G[i];           // G(i + r)
    T v;
    if (op == kSum) {
      const T sum = h + g;
      v = floor_mod(static_cast<int64_t>(i) - r - o, wa) == 0 ? sum * T(0.5) : sum;
    } else {
      v = apply(op, h, g);
    }
    y[i * sy] = v;
  }
}

// One pass along axis a (in may equal out); scale: multiply by fac after the sum.
template <class T>
void pass_axis(const Shape& s, int a, int32_t w, Op op, const T* in, T* out, bool scale, T fac) {
  std::vector<T> buf, tmp(static_cast<size_t>(s.n[a]));
  for_lines(s, a, [&](int64_t base) {
    vhgw_line(in + base, s.st[a], s.n[a], w, op, tmp.data(), 1, buf);
    for (int32_t i = 0; i < s.n[a]; ++i) out[base + i * s.st[a]] = scale ? tmp[i] * fac : tmp[i];
  });
}

// The separable cube (cpp/src/method/common.cpp:65-136): box axis 0 first, the last pass scaled by
// fl_T(1 / prod of the clamped sides); max / min the last axis first (exact: any order).
enum CubeOp { kCubeMax = 0, kCubeMin = 1, kCubeBox = 3 };
template <class T>
void cube(const Shape& s, int cop, const T* in, T* out, int32_t w) {
  if (cop == kCubeBox) {
    double prod = 1.0;
    for (int a = 0; a < s.D; ++a) prod *= static_cast<double>(2 * std::min((w - 1) / 2, s.n[a] - 1) + 1);
    const T fac = static_cast<T>(1.0 / prod);
    for (int a = 0; a < s.D; ++a) pass_axis(s, a, w, kSum, a == 0 ? in : out, out, a == s.D - 1, fac);
    return;
  }
  const Op op = cop == kCubeMax ? kMax : kMin;
  for (int a = s.D - 1; a >= 0; --a) pass_axis(s, a, w, op, a == s.D - 1 ? in : out, out, false, T(1));
}

// ------------------------------------------------------------------------------------------------
// Canonical reductions (cpp/src/field/common.cuh:134-259, reduce.hpp; twin.hpp:136-158): per
// last-axis row, sum v^2 in double over 32-sample chunks sequentially, then a pairwise tree over
// the chunk sums (padded with zeros to a power of two); per element the same tree over the rows.
inline double pairwise(std::vector<double>& s) {
  size_t m = 1;
  while (m < s.size()) m <<= 1;
  s.resize(std::max<size_t>(m, 1), 0.0);
  for (size_t d = 1; d < s.size(); d <<= 1)
    for (size_t i = 0; i + d < s.size(); i += 2 * d) s[i] = s[i] + s[i + d];
  return s[0];
}
template <class T>
double field_energy(const Shape& s, const T* f) {
  const int64_t nl = s.n[s.D - 1];
  std::vector<double> rows, chunks;
  for (int64_t q = 0; q < s.count / nl; ++q) {
    chunks.clear();
    for (int64_t c = 0; c < nl; c += 32) {
      double acc = 0.0;
      for (int64_t e = c; e < std::min(c + 32, nl); ++e) {
        const double d = static_cast<double>(f[q * nl + e]);
        acc = acc + d * d;
      }
      chunks.push_back(acc);
    }
    rows.push_back(pairwise(chunks));
  }
  return pairwise(rows);
}
// max |v| as the largest bit pattern of |v| (NaN above inf, as the device's integer maximum).
template <class T>
double field_amax(const Shape& s, const T* f) {
  uint64_t m = 0;
  for (int64_t i = 0; i < s.count; ++i) m = std::max(m, abs_bits(f[i]));
  return bits_d(m);
}

// ------------------------------------------------------------------------------------------------
// Exact selections (cpp/src/method/stats.cu:1-259): order statistics of doubles (-0 counts as +0,
// the order key's rule stats_common.cuh:52-62), the median as numpy.median, the quantile as
// numpy.quantile(method='linear') (stats_common.cuh:99-127).
inline double canon_zero(double v) { return v == 0.0 ? 0.0 : v; }
inline double kth(std::vector<double>& v, size_t k) {
  std::nth_element(v.begin(), v.begin() + static_cast<std::ptrdiff_t>(k), v.end());
  return v[k];
}
inline double select_median(std::vector<double>& v) {
  const size_t n = v.size();
  if (n == 0) return std::numeric_limits<double>::quiet_NaN();
  for (double& d : v) d = canon_zero(d);
  const double a = kth(v, (n - 1) / 2);
  if (n % 2) return a;
  const double b = kth(v, n / 2);
  return (a + b) / 2.0;
}
inline double select_quantile(std::vector<double>& v, double q) {
  const int64_t n = static_cast<int64_t>(v.size());
  if (n == 0) return std::numeric_limits<double>::quiet_NaN();
  for (double& d : v) d = canon_zero(d);
  const double idx = static_cast<double>(n - 1) * q;
  int64_t i, j;
  double gamma;
  if (idx >= static_cast<double>(n - 1)) {
    i = j = n - 1;
    gamma = idx + 1.0;
  } else if (idx < 0.0) {
    i = j = 0;
    gamma = idx;
  } else {
    const double f = std::floor(idx);
    i = static_cast<int64_t>(f);
    j = i + 1;
    gamma = idx - f;
  }
  const double a = kth(v, static_cast<size_t>(i));
  const double b = kth(v, static_cast<size_t>(j));
  const double diff = b - a;
  return gamma >= 0.5 ? b - diff * (1.0 - gamma) : a + diff * gamma;
}

// Forward differences along the axes of the mask `axes`, computed in T (stats.cu:42-55).
template <class T>
void diffs(const Shape& s, const T* r, int axes, std::vector<double>& out) {
  out.clear();
  for (int a = 0; a < s.D; ++a) {
    if (!((axes >> a) & 1)) continue;
    for_lines(s, a, [&](int64_t base) {
      for (int32_t i = 0; i + 1 < s.n[a]; ++i) {
        const T d = r[base + (i + 1) * s.st[a]] - r[base + i * s.st[a]];
        out.push_back(static_cast<double>(d));
      }
    });
  }
}
// sigma = (1.4826 MAD(d)) / sqrt 2 (stats.cu:262-334): pooled over the axes, or the largest per axis.
template <class T>
double noise_sigma(const Shape& s, const T* r, bool per_axis) {
  const double sqrt2 = std::sqrt(2.0);
  std::vector<double> d, dev;
  auto level = [&](int axes) {
    diffs(s, r, axes, d);
    const double c = select_median(d);  // reorders d: fine, the MAD needs the set only
    dev.resize(d.size());
    for (size_t i = 0; i < d.size(); ++i) dev[i] = std::fabs(d[i] - c);
    const double mad = select_median(dev);
    return (1.4826 * mad) / sqrt2;
  };
  if (!per_axis) return level((1 << s.D) - 1);
  double acc = 0.0;
  for (int a = 0; a < s.D; ++a) {
    if (s.n[a] < 2) continue;
    const double v = level(1 << a);
    if (v > acc) acc = v;
  }
  return acc;
}

// ------------------------------------------------------------------------------------------------
// The threshold field (cpp/src/method/stats_fields.cu:26-69,375-389): det = r - Box_Wd r,
// lev = sqrt(max(Box_Wl det^2, 0)), tau = max(theta max(lev, k sigma_n), tol[, floor2]).
// Returns false when Box det^2 or sigma_n is not finite (the element's non-finite flag).
template <class T>
bool tau_field(const Shape& s, const Ctl& c, const T* r, int32_t Wd, int32_t Wl, double sigma_n, double tol,
               const double* floor2, T* tau, T* tmp) {
  cube<T>(s, kCubeBox, r, tmp, Wd);
  for (int64_t i = 0; i < s.count; ++i) {
    const T det = r[i] - tmp[i];
    tau[i] = det * det;
  }
  cube<T>(s, kCubeBox, tau, tau, Wl);
  bool finite = true;
  const double f = c.sig_noise * sigma_n;
  const double fl = f > 0.0 ? f : 0.0;
  for (int64_t i = 0; i < s.count; ++i) {
    const T be = tau[i];
    if (!(std::isfinite(be) && std::isfinite(sigma_n))) finite = false;
    const T lev = std::sqrt(be > T(0) ? be : T(0));
    const double ld = static_cast<double>(lev);
    double t = c.sig_theta * (ld > fl ? ld : fl);
    t = t > tol ? t : tol;
    if (floor2) t = t > *floor2 ? t : *floor2;
    tau[i] = static_cast<T>(t);
  }
  return finite;
}

// Significant extrema (stats_fields.cu:1-14,162-331): along every line of every axis with n >= 3,
// the peaks of r and -r as scipy.signal.find_peaks (a strict rise, a plateau, a strict fall, never
// at the line ends, the plateau's peak at (left + right) // 2) whose topographic prominence
// (scipy.signal.peak_prominences) is >= tau at the peak: on each side, some sample before the
// first strictly higher one (or a NaN) has fl64(v - u) >= tau. Sets bit a of flags; counts ext[a].
template <class T>
inline int side_step(bool mx, T u, T v, T tau) {
  if (mx ? !(u <= v) : !(u >= v)) return 2;
  const double d = mx ? static_cast<double>(v) - static_cast<double>(u) : static_cast<double>(u) - static_cast<double>(v);
  return d >= static_cast<double>(tau) ? 1 : 0;
}
template <class T>
void sig_flags(const Shape& s, const T* r, const T* tau, uint8_t* flags, uint64_t* ext) {
  std::fill(flags, flags + s.count, uint8_t{0});
  for (int a = 0; a < s.D; ++a) {
    ext[a] = 0;
    const int32_t n = s.n[a];
    if (n < 3) continue;
    const int64_t st = s.st[a];
    for_lines(s, a, [&](int64_t base) {
      const T* x = r + base;
      for (int32_t i = 1; i <= n - 2; ++i) {
        const T cur = x[i * st], prev = x[(i - 1) * st];
        if (!(prev < cur || prev > cur)) continue;  // not a run start (equal or NaN)
        int32_t e = i;
        while (e + 1 < n && x[(e + 1) * st] == cur) ++e;
        if (e > n - 2) continue;
        const T nxt = x[(e + 1) * st];
        const bool is_max = prev < cur && nxt < cur;
        const bool is_min = prev > cur && nxt > cur;
        if (!is_max && !is_min) continue;
        const int32_t p = (i + e) / 2;
        const T tp = tau[base + p * st];
        auto side = [&](int dir) {
          for (int32_t k = p + dir; k >= 0 && k < n; k += dir) {
            const int st_ = side_step(is_max, x[k * st], cur, tp);
            if (st_) return st_ == 1;
          }
          return false;
        };
        if (side(-1) && side(1)) {
          flags[base + p * st] |= static_cast<uint8_t>(1u << a);
          ++ext[a];
        }
      }
    });
  }
}

// The raw window (stats_fields.cu:334-369,425-440): rho_a = Box_win(bit a of flags),
// dens = sqrt(sum rho_a^2) in T, om = 2 kappa / max(dens, tiny) where dens > 0, +inf elsewhere.
template <class T>
void raw_window(const Shape& s, const Ctl& c, const uint8_t* flags, int32_t win, T* om, T* tmp) {
  const T two_kappa = static_cast<T>(2.0 * c.kappa);
  const T tiny = static_cast<T>(c.tiny);
  for (int a = 0; a < s.D; ++a) {
    for (int64_t i = 0; i < s.count; ++i) tmp[i] = ((flags[i] >> a) & 1u) ? T(1) : T(0);
    cube<T>(s, kCubeBox, tmp, tmp, win);
    for (int64_t i = 0; i < s.count; ++i) {
      const T sq = tmp[i] * tmp[i];
      om[i] = a == 0 ? sq : om[i] + sq;
    }
  }
  for (int64_t i = 0; i < s.count; ++i) {
    const T dens = std::sqrt(om[i]);
    om[i] = dens > T(0) ? two_kappa / (dens > tiny ? dens : tiny) : std::numeric_limits<T>::infinity();
  }
}

// Face-connected components of mask (scipy.ndimage.label's default structure; the border does not
// connect), by breadth-first search: labels[v] = the smallest index of v's component, -1 outside
// (the reference's union-find gives the same minimum root: stats_scale.cu:1-83).
inline void ccl(const Shape& s, const uint8_t* mask, std::vector<int32_t>& labels) {
  labels.assign(static_cast<size_t>(s.count), -1);
  std::vector<int64_t> queue;
  for (int64_t v0 = 0; v0 < s.count; ++v0) {
    if (!mask[v0] || labels[v0] >= 0) continue;
    const int32_t lab = static_cast<int32_t>(v0);  // the first unvisited voxel in index order is its component's minimum
    labels[v0] = lab;
    queue.assign(1, v0);
    for (size_t h = 0; h < queue.size(); ++h) {
      const int64_t v = queue[h];
      int32_t x[3];
      coords_of(s, v, x);
      for (int a = 0; a < s.D; ++a) {
        for (int dir = -1; dir <= 1; dir += 2) {
          const int32_t y = x[a] + dir;
          if (y < 0 || y >= s.n[a]) continue;
          const int64_t u = v + dir * s.st[a];
          if (mask[u] && labels[u] < 0) {
            labels[u] = lab;
            queue.push_back(u);
          }
        }
      }
    }
  }
}

// The scale field (stats_scale.cu:85-167,222-246): the absence rule and L = exp(Box_W log(...)).
template <class T>
void scale_field(const Shape& s, const Ctl& c, int32_t w, int32_t W, double jump, const T* om, uint8_t* A, T* L,
                 T* fine) {
  const int32_t w_near = odd_window((2.0 * c.absent_reach) * static_cast<double>(W));
  const double wd = static_cast<double>(w);
  std::vector<uint8_t> aw(static_cast<size_t>(s.count)), a0(static_cast<size_t>(s.count));
  for (int64_t i = 0; i < s.count; ++i) {
    const double o = static_cast<double>(om[i]);
    a0[i] = o > c.reset * wd;
    aw[i] = o > c.absent_lo * wd;
    fine[i] = o < (c.absent_lo * wd) / jump ? T(1) : T(0);
  }
  cube<T>(s, kCubeMax, fine, fine, w_near);  // near
  std::vector<int32_t> labels;
  ccl(s, aw.data(), labels);
  std::vector<uint8_t> marks(static_cast<size_t>(s.count), 0);
  for (int64_t i = 0; i < s.count; ++i)
    if (aw[i] && fine[i] > T(0.5)) marks[labels[i]] = 1;
  for (int64_t i = 0; i < s.count; ++i) {
    const bool a = a0[i] && labels[i] >= 0 && marks[labels[i]];
    A[i] = a ? 1 : 0;
    double v;
    if (a) {
      v = c.absent;
    } else {
      const double o = static_cast<double>(om[i]);
      const double ow = o > c.reset * wd ? wd : o;
      const double lo = wd / c.c_lo;
      const double hi = c.c_hi * wd;
      v = ow > lo ? ow : lo;
      v = v < hi ? v : hi;
    }
    L[i] = std::log(static_cast<T>(v));
  }
  cube<T>(s, kCubeBox, L, L, W);
  for (int64_t i = 0; i < s.count; ++i) L[i] = std::exp(L[i]);
}

// The per-voxel floor for modes k >= 1 (stats_scale.cu:169-201,249-268).
template <class T>
void local_floor(const Shape& s, const Ctl& c, int32_t k, int32_t w, int32_t W, const T* Lp, const uint8_t* A, T* L,
                 T* chi) {
  for (int64_t i = 0; i < s.count; ++i) chi[i] = (A[i] & 1u) ? T(1) : T(0);
  cube<T>(s, kCubeBox, chi, chi, W);
  const T log_absent = static_cast<T>(std::log(c.absent));
  const bool grow = c.growth > 1.0 && k >= c.growth_from;
  const T cap = static_cast<T>(c.c_hi * static_cast<double>(w));
  for (int64_t i = 0; i < s.count; ++i) {
    const T lp = Lp[i];
    T fl = grow ? static_cast<T>(c.growth) * lp : lp + T(2);
    fl = fl < cap ? fl : cap;
    const T x = chi[i];
    const T e = (T(1) - x) * std::log(fl) + x * log_absent;
    const T f = std::exp(e);
    L[i] = L[i] >= f ? L[i] : f;
  }
}

// ------------------------------------------------------------------------------------------------
// G2's rung test and G3's coarse scale (cpp/src/method/stats_rungs.cu).
enum G2Outcome { kG2None = 0, kG2Accepted = 1, kG2Stop = 2 };
struct G2Result {
  int32_t outcome = kG2None, s = 0;
  int64_t n_p = 0;
  double p = 0.0, cov = 0.0;
};
template <class T>
G2Result g2_test(const Shape& s, const Ctl& c, int32_t w, int32_t srung, const uint8_t* ff, uint8_t* P, T* om,
                 T* tmp1) {
  const double sd = static_cast<double>(srung);
  const int32_t wb = odd_window(c.fine_beta * sd);
  const int32_t wcore = odd_window(c.fine_core * sd);
  const int32_t wr = odd_window((2.0 * c.absent_reach) * static_cast<double>(wb));
  raw_window<T>(s, c, ff, wb, om, tmp1);
  std::vector<double> vals;
  for (int64_t i = 0; i < s.count; ++i) {
    const bool in = static_cast<double>(om[i]) < c.fine_lo * sd;
    P[i] = in ? 1 : 0;
    tmp1[i] = in ? T(1) : T(0);
    if (in) vals.push_back(static_cast<double>(om[i]));
  }
  cube<T>(s, kCubeMin, tmp1, tmp1, wcore);
  bool core = false;
  for (int64_t i = 0; i < s.count; ++i) core = core || tmp1[i] > T(0.5);
  G2Result res;
  res.s = srung;
  res.n_p = static_cast<int64_t>(vals.size());
  res.p = select_quantile(vals, c.fine_pop);
  res.cov = static_cast<double>(res.n_p) / static_cast<double>(c.count);
  for (int64_t i = 0; i < s.count; ++i) tmp1[i] = (P[i] & 1u) ? T(1) : T(0);
  cube<T>(s, kCubeMax, tmp1, tmp1, wr);
  bool jump = false;
  for (int64_t i = 0; i < s.count; ++i)
    jump = jump || (static_cast<double>(om[i]) > c.reset * res.p && tmp1[i] > T(0.5));
  if (res.n_p == 0 || !core) res.outcome = kG2None;
  else if (c.fine_jump * res.p > static_cast<double>(w)) res.outcome = kG2Stop;
  else res.outcome = jump ? kG2Accepted : kG2None;
  return res;
}

template <class T>
double g3_coarse(const Shape& s, const Ctl& c, int32_t w, int32_t p_odd, const uint8_t* ff, const uint8_t* P, T* omc,
                 T* tmp1) {
  const int32_t Wc = odd_window(c.beta * static_cast<double>(w));
  const int32_t wf = odd_window(c.fine_beta * static_cast<double>(p_odd));
  const int32_t wr = odd_window((2.0 * c.absent_reach) * static_cast<double>(wf));
  raw_window<T>(s, c, ff, Wc, omc, tmp1);
  for (int64_t i = 0; i < s.count; ++i) tmp1[i] = (P[i] & 1u) ? T(1) : T(0);
  cube<T>(s, kCubeMax, tmp1, tmp1, wr);
  std::vector<double> vals;
  for (int64_t i = 0; i < s.count; ++i)
    if (tmp1[i] < T(0.5) && omc[i] < std::numeric_limits<T>::infinity()) vals.push_back(static_cast<double>(omc[i]));
  if (vals.empty()) return std::numeric_limits<double>::infinity();
  return select_median(vals);
}

// G2's candidate rungs and its confirmation machine (cpp/src/method/driver.cpp:59-116).
struct RungMachine {
  std::vector<int32_t> rungs;
  int32_t i = 0;
  bool pending = false, done = false, accepted = false;
  int32_t acc_index = -1;
  G2Result cand{};
  int32_t next_rung() const {
    if (done) return 0;
    if (i < static_cast<int32_t>(rungs.size())) return rungs[static_cast<size_t>(i)];
    return odd_window(static_cast<double>(rungs.back()) * kRungRatio + 1.0);
  }
  void feed(const G2Result& o) {
    if (pending) {
      if (o.outcome == kG2Accepted) {
        done = accepted = true;
        acc_index = i - 1;
        return;
      }
      if (o.outcome == kG2Stop) {
        done = true;
        return;
      }
      pending = false;
      ++i;
    } else {
      if (o.outcome == kG2Stop) {
        done = true;
        return;
      }
      if (o.outcome == kG2Accepted) {
        pending = true;
        cand = o;
        cand.s = rungs[static_cast<size_t>(i)];
      }
      ++i;
    }
    if (i >= static_cast<int32_t>(rungs.size()) && !pending) done = true;
  }
};
inline RungMachine make_rung_machine(int32_t w, int32_t w_prev, double jump) {
  RungMachine m;
  const int32_t s0 = std::max(3, w_prev + 2);
  for (int32_t j = 0;; ++j) {
    const int32_t sv = odd_window(static_cast<double>(s0) * std::pow(kRungRatio, static_cast<double>(j)));
    if (static_cast<double>(sv) * jump > static_cast<double>(w)) break;
    if (m.rungs.empty() || sv != m.rungs.back()) m.rungs.push_back(sv);
    if (static_cast<int32_t>(m.rungs.size()) > kMaxG2Rungs) throw std::length_error("too many G2 rungs");
  }
  m.done = m.rungs.empty();
  return m;
}

// ------------------------------------------------------------------------------------------------
// The orientation (cpp/src/method/orientation.cu:1-26,136-404; long-double twin
// cpp/tests/twin_orientation.hpp:1-120): X = the orthonormal N-D DCT-II of r - mean(r), the band
// X_b = G(rad) X (one-octave log-Gabor at f0 = 1/w), the Riesz components R1_a (sin basis on axis
// a), R2_aa, R2_ab (sin basis on a and b) of the band by the inverse transform, T = R1 R1^T +
// R2 R2^T and its Jacobi eigen-decomposition into n and c. The transforms are direct sums in
// double over angle tables cos(pi j / (2n)), j < 4n, evaluated in long double.
struct AxisTables {
  int32_t n = 0;
  std::vector<double> fwd, icos, isin;  // n x n: fwd[k][x]; icos/isin[x][k] (the c_k factor included)
};
inline AxisTables axis_tables(int32_t n) {
  AxisTables t;
  t.n = n;
  std::vector<long double> cs(4 * static_cast<size_t>(n)), sn(4 * static_cast<size_t>(n));
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int64_t j = 0; j < 4 * n; ++j) {
    cs[j] = std::cos(pi * static_cast<long double>(j) / static_cast<long double>(2 * n));
    sn[j] = std::sin(pi * static_cast<long double>(j) / static_cast<long double>(2 * n));
  }
  const size_t nn = static_cast<size_t>(n) * n;
  t.fwd.resize(nn);
  t.icos.resize(nn);
  t.isin.resize(nn);
  for (int64_t k = 0; k < n; ++k) {
    const long double ck = std::sqrt((k ? 2.0L : 1.0L) / static_cast<long double>(n));
    for (int64_t x = 0; x < n; ++x) {
      const int64_t q = (k * (2 * x + 1)) % (4 * n);
      t.fwd[k * n + x] = static_cast<double>(ck * cs[q]);
      t.icos[x * n + k] = static_cast<double>(ck * cs[q]);
      t.isin[x * n + k] = static_cast<double>(ck * sn[q]);
    }
  }
  return t;
}
// f <- M f along axis a (M row-major n x n), in double.
inline void axis_apply(const Shape& s, int a, const std::vector<double>& M, std::vector<double>& f) {
  const int32_t n = s.n[a];
  std::vector<double> line(static_cast<size_t>(n));
  for_lines(s, a, [&](int64_t base) {
    for (int32_t i = 0; i < n; ++i) line[i] = f[base + i * s.st[a]];
    for (int32_t i = 0; i < n; ++i) {
      const double* row = &M[static_cast<size_t>(i) * n];
      double acc = 0.0;
      for (int32_t j = 0; j < n; ++j) acc += row[j] * line[j];
      f[base + i * s.st[a]] = acc;
    }
  });
}
// The frequencies u_a = k_a (1 / (2 n_a)) and rad = sqrt(sum u_a^2) (orientation.cu:101-111).
inline double dct_freq(const Shape& s, const int32_t* k, double* u) {
  double r2 = 0.0;
  for (int a = 0; a < s.D; ++a) {
    u[a] = static_cast<double>(k[a]) * (1.0 / (2.0 * static_cast<double>(s.n[a])));
    r2 = r2 + u[a] * u[a];
  }
  return std::sqrt(r2);
}
inline double band_two_sigma2() {  // orientation.cu:67-79
  volatile double two = 2.0;
  const double l2 = std::log(static_cast<double>(two));
  const double sg = l2 / (2.0 * std::sqrt(2.0 * l2));
  return 2.0 * (sg * sg);
}
// The pair index of R2 / T (orientation.cu:365-369): 2D (0,0) (0,1) (1,1); 3D (0,0) (0,1) (0,2) (1,1) (1,2) (2,2).
inline int tensor_pair(int D, int a, int b) {
  if (a > b) std::swap(a, b);
  return D == 2 ? (a == 0 ? b : 2) : (a == 0 ? b : (a == 1 ? 2 + b : 5));
}

// One Jacobi rotation (cpp/src/method/element.cu:36-80, Rutishauser's form).
inline void jacobi_rotate(double (&A)[3][3], double (&V)[3][3], int p, int q, int r, int dim) {
  const double apq = A[p][q];
  if (apq == 0.0) return;
  const double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
  const double at = std::fabs(theta);
  const double t = std::copysign(1.0, theta) / (at + std::sqrt(theta * theta + 1.0));
  const double c = 1.0 / std::sqrt(t * t + 1.0);
  const double sn = t * c;
  const double tau = sn / (1.0 + c);
  const double tapq = t * apq;
  A[p][p] = A[p][p] - tapq;
  A[q][q] = A[q][q] + tapq;
  A[p][q] = A[q][p] = 0.0;
  if (r >= 0) {
    const double gg = A[r][p], hh = A[r][q];
    const double np_ = gg - sn * (hh + gg * tau);
    const double nq_ = hh + sn * (gg - hh * tau);
    A[r][p] = A[p][r] = np_;
    A[r][q] = A[q][r] = nq_;
  }
  for (int k = 0; k < dim; ++k) {
    const double gg = V[k][p], hh = V[k][q];
    V[k][p] = gg - sn * (hh + gg * tau);
    V[k][q] = hh + sn * (gg - hh * tau);
  }
}
// n (top eigenvector) and c = (mu1 - mu2) / (mu1 + mu2) of the symmetric D x D tensor t (pair order);
// false when an entry is not finite (element.cu:82-164).
inline bool eigen(int D, const double* t, double* nv, double* cv) {
  double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  bool finite = true;
  int k = 0;
  for (int a = 0; a < D; ++a)
    for (int c = a; c < D; ++c, ++k) {
      finite = finite && std::isfinite(t[k]);
      A[a][c] = A[c][a] = t[k];
    }
  if (D == 2) {
    jacobi_rotate(A, V, 0, 1, -1, 2);
  } else {
    double amax = 0.0;
    for (int a = 0; a < 3; ++a)
      for (int c = a; c < 3; ++c) amax = std::fmax(amax, std::fabs(A[a][c]));
    if (amax > 0.0 && amax < INFINITY) {
      const double sc = std::scalbn(1.0, -std::ilogb(amax));
      for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) A[a][c] = A[a][c] * sc;
    }
    for (int sweep = 0; sweep < 6; ++sweep) {
      jacobi_rotate(A, V, 0, 1, 2, 3);
      jacobi_rotate(A, V, 0, 2, 1, 3);
      jacobi_rotate(A, V, 1, 2, 0, 3);
    }
  }
  double mu[3];
  for (int a = 0; a < D; ++a) mu[a] = A[a][a];
  int top = 0;
  for (int a = 1; a < D; ++a) top = mu[a] > mu[top] ? a : top;
  double mu2 = -INFINITY;
  for (int a = 0; a < D; ++a) mu2 = (a != top && mu[a] > mu2) ? mu[a] : mu2;
  const double mu1 = mu[top];
  for (int a = 0; a < D; ++a) nv[a] = V[a][top];
  const double ssum = mu1 + mu2;
  double c = ssum > 0.0 ? (mu1 - mu2) / ssum : 0.0;
  *cv = c < 0.0 ? 0.0 : (c > 1.0 ? 1.0 : c);
  return finite;
}

struct OrientPlan {
  Shape s;
  AxisTables ax[3];
  double two_sigma2 = 0.0;
};
inline OrientPlan orient_plan(const Shape& s) {
  OrientPlan p;
  p.s = s;
  for (int a = 0; a < s.D; ++a) p.ax[a] = axis_tables(s.n[a]);
  p.two_sigma2 = band_two_sigma2();
  return p;
}

// The Riesz components of the band of window w into r1 (D fields) and r2 (D (D+1) / 2 fields, pair
// order), rounded to T where the reference's fields are T.
template <class T>
void riesz(const OrientPlan& P, const T* r, int32_t w, std::vector<std::vector<T>>& r1,
           std::vector<std::vector<T>>& r2) {
  const Shape& s = P.s;
  const int D = s.D;
  const int64_t N = s.count;
  double mean = 0.0;
  for (int64_t i = 0; i < N; ++i) mean += static_cast<double>(r[i]);
  mean = mean / static_cast<double>(N);
  std::vector<double> X(static_cast<size_t>(N));
  for (int64_t i = 0; i < N; ++i) X[i] = static_cast<double>(static_cast<T>(static_cast<double>(r[i]) - mean));
  for (int a = 0; a < D; ++a) axis_apply(s, a, P.ax[a].fwd, X);
  // the band, rounded to T (the reference keeps X and X_b in T)
  const double f0 = 1.0 / (1.0 * static_cast<double>(w));
  std::vector<double> rad(static_cast<size_t>(N)), uk(static_cast<size_t>(D * N));
  for (int64_t i = 0; i < N; ++i) {
    int32_t k[3];
    coords_of(s, i, k);
    double u[3];
    rad[i] = dct_freq(s, k, u);
    for (int a = 0; a < D; ++a) uk[a * N + i] = u[a];
    double G = 0.0;
    if (rad[i] > 0.0) {
      const double l = std::log(rad[i] / f0);
      G = std::exp(-(l * l) / P.two_sigma2);
    }
    const double xt = static_cast<double>(static_cast<T>(X[i]));
    X[i] = static_cast<double>(static_cast<T>(xt * G));
  }
  auto component = [&](int kind, int a, int b, std::vector<T>& out) {
    std::vector<double> y(static_cast<size_t>(N));
    for (int64_t i = 0; i < N; ++i) {
      double m = 0.0;
      if (rad[i] > 0.0) {
        const double inv = 1.0 / rad[i];
        const double ua = uk[a * N + i] * inv;
        if (kind == 0) m = ua;
        else if (kind == 1) m = (-ua) * ua;
        else m = ua * (uk[b * N + i] * inv);
      }
      y[i] = static_cast<double>(static_cast<T>(m * X[i]));
    }
    const int odd = kind == 0 ? (1 << a) : (kind == 1 ? 0 : ((1 << a) | (1 << b)));
    for (int ax = 0; ax < D; ++ax) axis_apply(s, ax, ((odd >> ax) & 1) ? P.ax[ax].isin : P.ax[ax].icos, y);
    out.resize(static_cast<size_t>(N));
    for (int64_t i = 0; i < N; ++i) out[i] = static_cast<T>(y[i]);
  };
  r1.resize(D);
  r2.resize(D * (D + 1) / 2);
  for (int a = 0; a < D; ++a) component(0, a, a, r1[a]);
  for (int a = 0; a < D; ++a)
    for (int b = a; b < D; ++b) component(a == b ? 1 : 2, a, b, r2[tensor_pair(D, a, b)]);
}

// n (D fields) and c of the element; false when the tensor has a non-finite entry.
template <class T>
bool orientation(const OrientPlan& P, const T* r, int32_t w, std::vector<std::vector<T>>& n, std::vector<T>& coh) {
  const Shape& s = P.s;
  const int D = s.D;
  std::vector<std::vector<T>> r1, r2;
  riesz<T>(P, r, w, r1, r2);
  n.assign(D, std::vector<T>(static_cast<size_t>(s.count)));
  coh.resize(static_cast<size_t>(s.count));
  bool finite = true;
  const int K = D * (D + 1) / 2;
  for (int64_t i = 0; i < s.count; ++i) {
    T R2[6], out[6];
    for (int p = 0; p < K; ++p) R2[p] = r2[p][i];
    for (int a = 0; a < D; ++a)
      for (int c = a; c < D; ++c) {
        T v = r1[a][i] * r1[c][i];
        for (int k = 0; k < D; ++k) v = v + R2[tensor_pair(D, a, k)] * R2[tensor_pair(D, k, c)];
        out[tensor_pair(D, a, c)] = v;
      }
    double t[6], nv[3], cv;
    for (int p = 0; p < K; ++p) t[p] = static_cast<double>(out[p]);
    finite = eigen(D, t, nv, &cv) && finite;
    coh[i] = static_cast<T>(cv);
    for (int a = 0; a < D; ++a) n[a][i] = static_cast<T>(nv[a]);
  }
  return finite;
}

// ------------------------------------------------------------------------------------------------
// The monogenic signal (cpp/src/method/orientation.cu:1-26,86-128,420-493,789-831;
// monogenic.cpp:61-158): f = x 2^-k with k = ilogb(max |x|) (0 for a zero or non-finite max),
// X = the orthonormal DCT-II of f - mean(f) (rounded to T), no band, and the components of the whole
// spectrum by the inverse transform with the sin basis on their odd axes, each multiplier 0 at the DC:
//   R1_a u_a / rad (sin a); R2_aa -(u_a / rad)^2; R2_ab (u_a / rad)(u_b / rad) (sin a, b);
//   D_a -u_a (sin a); D_a R_a u_a (u_a / rad); D_b R_a -u_a (u_b / rad) (sin a, b), a < b.
// D >= 2: n, c from T = R1 R1^T + R2 R2^T (n with the canonical sign: its first component of
// magnitude above the dtype's default tie tolerance positive); n = 1 in 1-D. Per voxel in double:
// q = n . R1, A^2 = f^2 + |R1|^2, D_n f = n . D f, D_n q = sum_{a <= b} m_ab D_b R_a (m_aa = n_a^2,
// m_ab = 2 n_a n_b); where A^2 > tiny: phase = atan2(q, f) (a phase that rounds to -pi in T is pi),
// frequency = (f D_n q - q D_n f) / A^2, else both 0; amplitude = sqrt(A^2) 2^k rounded once.
// out: amplitude, phase, frequency, then (D >= 2) the D direction planes and the coherence, count
// samples each.
template <class T>
void monogenic(const OrientPlan& P, const T* x, T* out) {
  const Shape& s = P.s;
  const int D = s.D;
  const int64_t N = s.count;
  double amax = 0.0;
  for (int64_t i = 0; i < N; ++i) amax = std::fmax(amax, std::fabs(static_cast<double>(x[i])));
  bool finite = true;
  for (int64_t i = 0; i < N; ++i) finite = finite && std::isfinite(static_cast<double>(x[i]));
  const int k = (amax > 0.0 && finite) ? std::ilogb(amax) : 0;
  std::vector<T> f(static_cast<size_t>(N));
  for (int64_t i = 0; i < N; ++i) f[i] = static_cast<T>(std::ldexp(static_cast<double>(x[i]), -k));
  double mean = 0.0;
  for (int64_t i = 0; i < N; ++i) mean += static_cast<double>(f[i]);
  mean = mean / static_cast<double>(N);
  std::vector<double> X(static_cast<size_t>(N));
  for (int64_t i = 0; i < N; ++i) X[i] = static_cast<double>(static_cast<T>(static_cast<double>(f[i]) - mean));
  for (int a = 0; a < D; ++a) axis_apply(s, a, P.ax[a].fwd, X);
  std::vector<double> rad(static_cast<size_t>(N)), uk(static_cast<size_t>(D * N));
  for (int64_t i = 0; i < N; ++i) {
    int32_t kk[3];
    coords_of(s, i, kk);
    double u[3];
    rad[i] = dct_freq(s, kk, u);
    for (int a = 0; a < D; ++a) uk[a * N + i] = u[a];
    X[i] = static_cast<double>(static_cast<T>(X[i]));
  }
  auto component = [&](int kind, int a, int b) {
    std::vector<double> y(static_cast<size_t>(N));
    for (int64_t i = 0; i < N; ++i) {
      double m = 0.0;
      if (rad[i] > 0.0) {
        const double inv = 1.0 / rad[i];
        const double ua = uk[a * N + i] * inv;
        if (kind == 0) m = ua;
        else if (kind == 1) m = (-ua) * ua;
        else if (kind == 2) m = ua * (uk[b * N + i] * inv);
        else if (kind == 3) m = -uk[a * N + i];
        else m = a == b ? uk[a * N + i] * ua : -(uk[a * N + i] * (uk[b * N + i] * inv));
      }
      y[i] = static_cast<double>(static_cast<T>(m * X[i]));
    }
    int odd = 0;
    if (kind == 0 || kind == 3) odd = 1 << a;
    else if (kind == 2 || (kind == 4 && a != b)) odd = (1 << a) | (1 << b);
    for (int ax = 0; ax < D; ++ax) axis_apply(s, ax, ((odd >> ax) & 1) ? P.ax[ax].isin : P.ax[ax].icos, y);
    std::vector<T> o(static_cast<size_t>(N));
    for (int64_t i = 0; i < N; ++i) o[i] = static_cast<T>(y[i]);
    return o;
  };
  const int K = D * (D + 1) / 2;
  std::vector<std::vector<T>> r1(D), df(D), dr(K);
  for (int a = 0; a < D; ++a) r1[a] = component(0, a, a);
  for (int a = 0; a < D; ++a) df[a] = component(3, a, a);
  for (int a = 0; a < D; ++a)
    for (int b = a; b < D; ++b) dr[tensor_pair(D == 1 ? 2 : D, a, b)] = component(4, a, b);
  std::vector<std::vector<T>> nv(D, std::vector<T>(static_cast<size_t>(N), T(1)));
  std::vector<T> coh(static_cast<size_t>(N), T(0));
  if (D >= 2) {
    std::vector<std::vector<T>> r2(K);
    for (int a = 0; a < D; ++a)
      for (int b = a; b < D; ++b) r2[tensor_pair(D, a, b)] = component(a == b ? 1 : 2, a, b);
    for (int64_t i = 0; i < N; ++i) {
      T out6[6];
      for (int a = 0; a < D; ++a)
        for (int c = a; c < D; ++c) {
          T v = r1[a][i] * r1[c][i];
          for (int q = 0; q < D; ++q) v = v + r2[tensor_pair(D, a, q)][i] * r2[tensor_pair(D, q, c)][i];
          out6[tensor_pair(D, a, c)] = v;
        }
      double t[6], n3[3], cv;
      for (int p = 0; p < K; ++p) t[p] = static_cast<double>(out6[p]);
      eigen(D, t, n3, &cv);
      coh[i] = static_cast<T>(cv);
      for (int a = 0; a < D; ++a) nv[a][i] = static_cast<T>(n3[a]);
    }
  }
  const T zero = static_cast<T>(std::is_same_v<T, float> ? std::ldexp(1.0, -14) : std::ldexp(1.0, -30));
  const T pi_t = static_cast<T>(3.141592653589793);
  for (int64_t i = 0; i < N; ++i) {
    T nt[3];
    int sgn = 0;
    for (int a = 0; a < D; ++a) {
      nt[a] = nv[a][i];
      if (D >= 2 && sgn == 0 && std::fabs(nt[a]) > zero) sgn = nt[a] > T(0) ? 1 : -1;
    }
    for (int a = 0; a < D; ++a) nt[a] = sgn < 0 ? -nt[a] : nt[a];
    double n[3];
    for (int a = 0; a < D; ++a) n[a] = static_cast<double>(nt[a]);
    const double fv = static_cast<double>(f[i]);
    double q = 0.0, a2 = fv * fv, dnf = 0.0, dnq = 0.0;
    for (int a = 0; a < D; ++a) {
      const double r = static_cast<double>(r1[a][i]);
      q = q + n[a] * r;
      a2 = a2 + r * r;
      dnf = dnf + n[a] * static_cast<double>(df[a][i]);
    }
    for (int a = 0; a < D; ++a)
      for (int c = a; c < D; ++c) {
        const double m = a == c ? n[a] * n[a] : 2.0 * (n[a] * n[c]);
        dnq = dnq + m * static_cast<double>(dr[tensor_pair(D == 1 ? 2 : D, a, c)][i]);
      }
    double phase = 0.0, freq = 0.0;
    const double tiny = std::is_same_v<T, float> ? static_cast<double>(FLT_MIN) : 1e-300;
    if (a2 > tiny) {
      phase = std::atan2(q, fv);
      freq = (fv * dnq - q * dnf) / a2;
    }
    out[i] = static_cast<T>(std::ldexp(std::sqrt(a2), k));
    const T pt = static_cast<T>(phase);
    out[N + i] = pt == -pi_t ? pi_t : pt;
    out[2 * N + i] = static_cast<T>(freq);
    if (D >= 2) {
      for (int a = 0; a < D; ++a) out[(3 + a) * N + i] = nt[a];
      out[(3 + D) * N + i] = coh[i];
    }
  }
}

// ------------------------------------------------------------------------------------------------
// The element (cpp/src/method/element.hpp:1-29, element.cuh:53-69, element.cu:166-303).
template <class T>
struct DE {
  T d, e;
};
template <class T>
inline DE<T> element_de(T L, T c, T q, T span_m1) {
  const T f = (T(1) - c) + c / q;
  DE<T> r;
  r.d = L * f;
  const T len = L * (T(1) + span_m1 * c);
  const T diff = len - r.d;
  r.e = T(0.5) * (diff > T(0) ? diff : T(0));
  return r;
}

// The rung ladder (element.cu:343-370): in double with the C library's log and pow.
inline std::vector<int32_t> rung_ladder(double d_min, double d_max, int32_t w, int32_t n_max) {
  volatile double qv = kRungRatio;
  const double q = qv;
  const double lq = std::log(q);
  const double wd = static_cast<double>(w);
  const double jl = std::floor(std::log((d_min > 1.0 ? d_min : 1.0) / wd) / lq + 1e-9);
  const double jh = std::ceil(std::log((d_max > 1.0 ? d_max : 1.0) / wd) / lq - 1e-9);
  if (!(jh - jl <= 4096.0)) throw std::length_error("ladder spans too many rungs");
  const int64_t cap = 2 * static_cast<int64_t>(n_max) + 1;
  std::set<int64_t> sides;
  for (int64_t j = static_cast<int64_t>(jl); j <= static_cast<int64_t>(jh); ++j) {
    const double v = wd * std::pow(q, static_cast<double>(j));
    const int64_t side = v >= static_cast<double>(cap) + 1.0 ? cap : std::min<int64_t>(odd_window(v), cap);
    sides.insert(side);
  }
  if (sides.size() > static_cast<size_t>(kMaxRungs)) throw std::length_error("ladder exceeds 32 rungs");
  return std::vector<int32_t>(sides.begin(), sides.end());
}

// The element of one mode as the operators read it.
template <class T>
struct Element {
  int K = 0;                      // rungs
  int32_t rung_w[kMaxRungs] = {};
  uint32_t used = 0, pts = 0;     // per-rung bits
  std::vector<uint8_t> lo;
  std::vector<T> t;
  std::vector<std::vector<T>> o;  // D components (D >= 2)
};

// D-linear interpolation of f at y with half-sample periodic reflection (element_op.cuh:45-75).
template <class T>
T interp(const Shape& s, const T* f, const double* y) {
  const int D = s.D;
  int64_t off0[3], off1[3];
  T w0[3], w1[3];
  for (int a = 0; a < D; ++a) {
    const double fl = std::floor(y[a]);
    const double tt = y[a] - fl;
    const int64_t i0 = static_cast<int64_t>(fl);
    const T t1 = static_cast<T>(tt);
    w1[a] = t1;
    w0[a] = T(1) - t1;
    off0[a] = reflect_any(i0, s.n[a]) * s.st[a];
    off1[a] = reflect_any(i0 + 1, s.n[a]) * s.st[a];
  }
  T acc = T(0);
  for (int corner = 0; corner < (1 << D); ++corner) {
    int64_t off = 0;
    for (int a = 0; a < D; ++a) off += ((corner >> (D - 1 - a)) & 1) ? off1[a] : off0[a];
    T term = f[off];
    for (int a = 0; a < D; ++a) term = term * (((corner >> (D - 1 - a)) & 1) ? w1[a] : w0[a]);
    acc = acc + term;
  }
  return acc;
}

// The operators Dil_B / Ero_B / Avg_B (element_op.cuh:77-263): per rung j in ascending order, the
// rung's cube and the gather with the hat weight. kMaxMin writes Dil into out0 and Ero into out1.
enum ElemOp { kOpMax = 0, kOpMaxMin = 2, kOpBox = 3 };
template <class T>
void element_op(const Shape& s, int op, const Element<T>& el, const T* in, T* out0, T* out1) {
  const int D = s.D;
  std::vector<T> f0(static_cast<size_t>(s.count)), f1;
  if (op == kOpMaxMin) f1.resize(static_cast<size_t>(s.count));
  for (int j = 0; j < el.K; ++j) {
    if (!((el.used >> j) & 1u)) continue;
    const int32_t wj = el.rung_w[j];
    if (op == kOpBox) {
      cube<T>(s, kCubeBox, in, f0.data(), wj);
    } else {
      cube<T>(s, kCubeMax, in, f0.data(), wj);
      if (op == kOpMaxMin) cube<T>(s, kCubeMin, in, f1.data(), wj);
    }
    const bool pts = D >= 2 && ((el.pts >> j) & 1u);
    for (int64_t i = 0; i < s.count; ++i) {
      const int l = el.lo[i];
      const T tt = el.t[i];
      const T wl = T(1) - tt;
      T wgt;
      bool add;
      if (l == j && wl > T(0)) {
        wgt = wl;
        add = false;
      } else if (l + 1 == j && tt > T(0)) {
        wgt = tt;
        add = wl > T(0);
      } else {
        continue;
      }
      T acc0 = f0[i];
      T acc1 = op == kOpMaxMin ? f1[i] : T(0);
      if (pts) {
        int32_t x[3];
        coords_of(s, i, x);
        double ov[3];
        for (int a = 0; a < D; ++a) ov[a] = static_cast<double>(el.o[a][i]);
        for (int k = 1; k <= 2; ++k) {
          const double sc = static_cast<double>(k) / 2.0;
          for (int sg = 0; sg < 2; ++sg) {
            double y[3];
            for (int a = 0; a < D; ++a) {
              const double off = sc * ov[a];
              const double xa = static_cast<double>(x[a]);
              y[a] = sg == 0 ? xa + off : xa - off;
            }
            if (op == kOpMaxMin) {
              acc0 = max_c(acc0, interp(s, f0.data(), y));
              acc1 = min_c(acc1, interp(s, f1.data(), y));
            } else if (op == kOpMax) {
              acc0 = max_c(acc0, interp(s, f0.data(), y));
            } else {
              acc0 = acc0 + interp(s, f0.data(), y);
            }
          }
        }
        if (op == kOpBox) acc0 = acc0 / T(5);
      }
      const T v0 = wgt * acc0;
      out0[i] = add ? out0[i] + v0 : v0;
      if (op == kOpMaxMin) {
        const T v1 = wgt * acc1;
        out1[i] = add ? out1[i] + v1 : v1;
      }
    }
  }
}

// The envelope mean (element_op.cuh:266-282): u = Dil h, l = Ero h, m = Avg(0.5 (u + l)).
template <class T>
void envelope_mean(const Shape& s, const Element<T>& el, const T* h, T* m, T* u, T* l) {
  element_op<T>(s, kOpMaxMin, el, h, u, l);
  std::vector<T> mid(static_cast<size_t>(s.count));
  for (int64_t i = 0; i < s.count; ++i) mid[i] = T(0.5) * (u[i] + l[i]);
  element_op<T>(s, kOpBox, el, mid.data(), m, nullptr);
}

// ------------------------------------------------------------------------------------------------
// One element's decomposition (cpp/src/method/driver.cpp:393-700 per element; mode.cu, sift.cu).
enum { kErrNone = 0, kErrNonFinite = 1, kErrNonFiniteResult = 2, kErrInternal = 255 };

int32_t mode_bound(const Ctl& c) {  // driver.cpp:128-140
  int32_t w_prev = 1, count = 0;
  for (int32_t k = 0;; ++k) {
    int64_t w = std::max(3, w_prev + 2);
    if (c.growth > 1.0 && k >= c.growth_from) w = std::max<int64_t>(w, odd_window(c.growth * static_cast<double>(w_prev)));
    if (w > c.n_max) break;
    ++count;
    w_prev = static_cast<int32_t>(w);
  }
  return count;
}

struct ElemOut {
  int32_t n_imfs = 0, error = 0;
  std::vector<int32_t> window, sifts;
  std::vector<double> spacing, energy;
};

// The frozen decisions of one element's decomposition, one record per mode, for the reference of
// the adjoint (the sift map with windows, rungs, selections and labels frozen): the residual r at
// the mode's start, the scale field L, the lower rung lo of every voxel, and per mode the integers
// [w, W, K, used, pts, rung_w[0..31], sifts run after the first, accepted bits (bit j: sift j),
// kexp] and the doubles [tol, r_inf, thr_j (j = 1 .. sifts - 1, at index 1 + j)].
constexpr int kTapeInts = 40;
constexpr int kTapeDbls = 34;
struct TapeSink {
  int32_t modes = 0;  // records available
  void* r = nullptr;
  void* L = nullptr;
  uint8_t* lo = nullptr;
  int32_t* ints = nullptr;
  double* dbls = nullptr;
};

template <class T>
ElemOut decompose_one(const Shape& s, const Ctl& c, const T* x, int32_t max_imf, double tau_rel, bool check_finite,
                      int32_t cap, T* rows /* cap * count */, const OrientPlan* plan, const TapeSink* tape = nullptr) {
  ElemOut out;
  const int64_t N = s.count;
  const int D = s.D;
  const double tau = static_cast<double>(static_cast<T>(tau_rel));
  std::vector<T> r(x, x + N);
  int64_t nonfinite = 0;
  for (int64_t i = 0; i < N; ++i) nonfinite += std::isfinite(static_cast<double>(r[i])) ? 0 : 1;
  // prescale by 2^-ilogb(max|x|) (mode.cu:43-58,271-282)
  const double amax0 = field_amax(s, r.data());
  const int32_t kexp = (amax0 > 0.0 && std::isfinite(amax0)) ? std::ilogb(amax0) : 0;
  if (kexp != 0) {
    for (int64_t i = 0; i < N; ++i) {
      if constexpr (std::is_same_v<T, float>) r[i] = static_cast<float>(std::ldexp(static_cast<double>(r[i]), -kexp));
      else r[i] = std::ldexp(r[i], -kexp);
    }
  }
  double amax = field_amax(s, r.data());
  double tol = tie_tolerance<T>(tau, amax);
  enum { kActive, kDone, kError } status = kActive;
  int32_t err = kErrNone;
  if (nonfinite != 0 && check_finite) {
    status = kError;
    err = kErrNonFinite;
  }
  const double sigma_x = status == kActive ? noise_sigma(s, r.data(), true) : 0.0;
  int32_t n_imfs = 0;
  int32_t w_prev = 1, k = 0;
  std::vector<T> L(N), Lp(N), h(N), tau_f(N), tmp(N), om(N), t1(N);
  std::vector<T> m(N), u(N), l(N), e0(N), e(N), q(N), Mn(N), ud(N), ld(N), hn(N), a(N), kap(N), gk(N), mn(N);
  std::vector<uint8_t> flags(N), ff(N), Pb[2] = {std::vector<uint8_t>(N), std::vector<uint8_t>(N)}, Pacc(N), A(N);
  const auto fail = [&](int32_t code) {
    status = kError;
    err = code;
  };
  auto sift_rows = [&](int64_t row) { return rows + row * N; };
  out.window.assign(static_cast<size_t>(cap - 1), 0);
  out.sifts.assign(static_cast<size_t>(cap - 1), 0);
  out.spacing.assign(static_cast<size_t>(cap - 1), 0.0);
  out.energy.assign(static_cast<size_t>(cap - 1), 0.0);

  while (status == kActive) {
    // step 0: mode start (mode.cu:95-136); sigma_n
    const double r_inf = amax;
    const double floor2 = c.fine_theta * sigma_x;
    uint64_t ext[3] = {0, 0, 0};
    auto tau_windows = [&](int32_t wt, int32_t* wd, int32_t* wl) {
      *wd = odd_window(c.sig_rho * static_cast<double>(wt));
      *wl = odd_window(std::max(c.beta, c.sig_rho) * static_cast<double>(wt));
    };
    int32_t Wd, Wl;
    tau_windows(std::max(3, w_prev + 2), &Wd, &Wl);
    const double sigma_n = noise_sigma(s, r.data(), false);
    int32_t w = 0;
    double spacing = 0.0;
    bool stop = false;
    // steps 1-2: the window passes (mode.cu:138-187)
    for (int32_t pass = 0; pass < c.sig_passes && !stop; ++pass) {
      if (!tau_field<T>(s, c, r.data(), Wd, Wl, sigma_n, tol, nullptr, tau_f.data(), tmp.data())) {
        fail(kErrNonFiniteResult);
        stop = true;
        break;
      }
      sig_flags<T>(s, r.data(), tau_f.data(), flags.data(), ext);
      const uint64_t NN = static_cast<uint64_t>(N);
      uint64_t S2 = 0;
      bool enough = false;
      for (int ax = 0; ax < D; ++ax) {
        S2 += ext[ax] * ext[ax];
        if (ext[ax] * static_cast<uint64_t>(s.n[ax]) >= 3 * NN) enough = true;
      }
      if (!enough) {
        status = kDone;
        stop = true;
        break;
      }
      const double sp = static_cast<double>(NN) / std::sqrt(static_cast<double>(S2));
      const double v = (2.0 * c.kappa) * sp;
      if (!(v < 1073741824.0)) {
        status = kDone;
        stop = true;
        break;
      }
      int32_t wv = 2 * static_cast<int32_t>(std::floor(std::max(v, 1.0) / 2.0)) + 1;
      wv = std::max(wv, 3);
      wv = std::max(wv, w_prev + 2);
      if (wv > c.n_max) {
        status = kDone;
        stop = true;
        break;
      }
      spacing = sp;
      tau_windows(wv, &Wd, &Wl);
      w = wv;
    }
    if (stop) break;
    // step 3: G2 (driver.cpp:445-509)
    bool fine = false;
    int32_t p_odd = 0;
    double jump = c.absent_jump, g3 = 1.0;
    const uint8_t* ffp = flags.data();
    const int32_t s0 = std::max(3, w_prev + 2);
    if (c.fine_jump * static_cast<double>(s0) <= static_cast<double>(w)) {
      RungMachine mach = make_rung_machine(w, w_prev, c.fine_jump);
      if (!mach.done) {
        bool nonfin = false;
        if (c.fine_theta > 0) {
          nonfin = !tau_field<T>(s, c, r.data(), Wd, Wl, sigma_n, tol, &floor2, tau_f.data(), tmp.data());
          uint64_t ext2[3];
          sig_flags<T>(s, r.data(), tau_f.data(), ff.data(), ext2);
          ffp = ff.data();
        }
        for (int32_t step = 0; !mach.done; ++step) {
          if (step > kMaxG2Rungs + 1) throw std::logic_error("G2 loop did not end");
          const int32_t sr = mach.next_rung();
          const G2Result res = g2_test<T>(s, c, w, sr, ffp, Pb[step & 1].data(), om.data(), t1.data());
          mach.feed(res);
        }
        if (mach.accepted) {
          fine = true;
          p_odd = odd_window(mach.cand.p);
          Pacc = Pb[mach.acc_index & 1];
        }
        if (nonfin) {
          fail(kErrNonFiniteResult);
          break;
        }
      }
    }
    double wc = std::numeric_limits<double>::infinity();
    if (fine) wc = g3_coarse<T>(s, c, w, p_odd, ffp, Pacc.data(), om.data(), t1.data());
    // step 4: the window (mode.cu:189-223)
    if (fine) {
      w = std::max(std::max(p_odd, w_prev + 2), 3);
      jump = 1.0;
    }
    if (c.growth > 1.0 && k >= c.growth_from) {
      w = std::max(w, odd_window(c.growth * static_cast<double>(w_prev)));
      if (w > c.n_max) {
        status = kDone;
        break;
      }
    }
    const int32_t W = odd_window(c.beta * static_cast<double>(w));
    if (fine && std::isfinite(wc)) {
      const double qv = std::sqrt(wc / static_cast<double>(p_odd));
      g3 = std::min((c.split * qv) / 2.0, 1.0);
    }
    // step 5: the scale field, the floor, G3's scaling
    raw_window<T>(s, c, flags.data(), W, om.data(), tmp.data());
    scale_field<T>(s, c, w, W, jump, om.data(), A.data(), L.data(), tmp.data());
    if (k >= 1) local_floor<T>(s, c, k, w, W, Lp.data(), A.data(), L.data(), tmp.data());
    if (fine) {
      for (int64_t i = 0; i < N; ++i) {
        const T v = static_cast<T>(g3) * L[i];
        L[i] = v >= T(3) ? v : T(3);
      }
    }
    // step 6: the orientation and the element build (element.cu:166-216)
    std::vector<std::vector<T>> nvec;
    std::vector<T> coh(static_cast<size_t>(N), T(0));
    if (D >= 2) {
      if (!orientation<T>(*plan, r.data(), w, nvec, coh)) {
        fail(kErrNonFiniteResult);
        break;
      }
    }
    Element<T> el;
    const T qT = static_cast<T>(c.q), span_m1 = static_cast<T>(c.span - 1.0);
    double d_min = INFINITY, d_max = 0.0;
    if (D >= 2) el.o.assign(D, std::vector<T>(static_cast<size_t>(N)));
    for (int64_t i = 0; i < N; ++i) {
      const DE<T> de = element_de<T>(L[i], D >= 2 ? coh[i] : T(0), qT, span_m1);
      if (D >= 2)
        for (int ax = 0; ax < D; ++ax) el.o[ax][i] = de.e * nvec[ax][i];
      d_min = std::min(d_min, static_cast<double>(de.d));
      d_max = std::max(d_max, static_cast<double>(de.d));
    }
    // step 7: the ladder and the hat weights (element.cu:218-292)
    const std::vector<int32_t> ladder = rung_ladder(d_min, d_max, w, c.n_max);
    el.K = static_cast<int>(ladder.size());
    for (int j = 0; j < el.K; ++j) el.rung_w[j] = ladder[j];
    double lw[kMaxRungs];
    for (int j = 0; j < el.K; ++j) lw[j] = std::log(static_cast<double>(el.rung_w[j]));
    el.lo.resize(static_cast<size_t>(N));
    el.t.resize(static_cast<size_t>(N));
    const double w_lo = static_cast<double>(el.rung_w[0]), w_hi = static_cast<double>(el.rung_w[el.K - 1]);
    for (int64_t i = 0; i < N; ++i) {
      const DE<T> de = element_de<T>(L[i], D >= 2 ? coh[i] : T(0), qT, span_m1);
      int idx = 0;
      T tt = T(0);
      if (el.K > 1) {
        double dc = static_cast<double>(de.d);
        dc = dc > w_lo ? dc : w_lo;
        dc = dc < w_hi ? dc : w_hi;
        const double lg = std::log(dc);
        int cnt = 0;
        for (int j = 0; j < el.K; ++j) cnt += lw[j] <= lg ? 1 : 0;
        idx = std::min(std::max(cnt - 1, 0), el.K - 2);
        tt = static_cast<T>((lg - lw[idx]) / (lw[idx + 1] - lw[idx]));
      }
      el.lo[i] = static_cast<uint8_t>(idx);
      el.t[i] = tt;
      const T wl = T(1) - tt;
      uint32_t bits = 0;
      if (wl > T(0)) bits |= 1u << idx;
      if (tt > T(0)) bits |= 1u << (idx + 1);
      el.used |= bits;
      if (de.e > T(0.25)) el.pts |= bits;
    }
    int32_t* tape_i = nullptr;
    double* tape_d = nullptr;
    if (tape != nullptr && n_imfs < tape->modes) {
      const size_t off = static_cast<size_t>(n_imfs) * N;
      std::copy(r.begin(), r.end(), static_cast<T*>(tape->r) + off);
      std::copy(L.begin(), L.end(), static_cast<T*>(tape->L) + off);
      std::copy(el.lo.begin(), el.lo.end(), tape->lo + off);
      tape_i = tape->ints + static_cast<size_t>(n_imfs) * kTapeInts;
      tape_d = tape->dbls + static_cast<size_t>(n_imfs) * kTapeDbls;
      std::fill(tape_i, tape_i + kTapeInts, 0);
      std::fill(tape_d, tape_d + kTapeDbls, 0.0);
      tape_i[0] = w;
      tape_i[1] = W;
      tape_i[2] = el.K;
      tape_i[3] = static_cast<int32_t>(el.used);
      tape_i[4] = static_cast<int32_t>(el.pts);
      for (int j = 0; j < el.K; ++j) tape_i[5 + j] = el.rung_w[j];
      tape_i[39] = kexp;
      tape_d[0] = tol;
      tape_d[1] = r_inf;
    }
    // step 8: sift 1 and the continuation weight (driver.cpp:606-638; sift.cu:55-85)
    envelope_mean<T>(s, el, r.data(), m.data(), u.data(), l.data());
    for (int64_t i = 0; i < N; ++i) h[i] = r[i] - m[i];
    int32_t n_sifts = 1;
    bool sifting = true;
    if (c.sifts > 1) {
      std::vector<T> R(N), U(N), U2(N), gw(N);
      for (int64_t i = 0; i < N; ++i) {
        const T dv = u[i] - l[i];
        R[i] = dv >= T(0) ? dv : T(0);
      }
      const T* Ucur = R.data();
      std::vector<T>* pp[2] = {&U, &U2};
      for (int32_t it = 0; it < c.u; ++it) {
        element_op<T>(s, kOpMax, el, Ucur, pp[it & 1]->data(), nullptr);
        Ucur = pp[it & 1]->data();
      }
      const T t0 = static_cast<T>(c.t0);
      const double dd = c.t1 - c.t0;
      const T den = static_cast<T>(dd >= 1e-12 ? dd : 1e-12);
      for (int64_t i = 0; i < N; ++i) {
        const T Uv = Ucur[i];
        T rho = T(0);
        if (static_cast<double>(Uv) > tol) rho = clip<T>(R[i] / Uv, T(0), T(1));
        gw[i] = clip<T>((rho - t0) / den, T(0), T(1));
      }
      const T* cur = gw.data();
      std::vector<T> tmpa(N);
      for (int32_t it = 0; it < c.s; ++it) {
        T* dst = it + 1 < c.s ? (cur == tmpa.data() ? gw.data() : tmpa.data()) : a.data();
        element_op<T>(s, kOpBox, el, cur, dst, nullptr);
        cur = dst;
      }
      for (int64_t i = 0; i < N; ++i) a[i] = clip<T>(a[i], T(0), T(1));
    }
    // step 9: sifts 2..S with G1 (driver.cpp:641-680; sift.cu:88-232)
    for (int32_t j = 1; j < c.sifts && sifting; ++j) {
      envelope_mean<T>(s, el, h.data(), m.data(), ud.data(), ld.data());
      for (int64_t i = 0; i < N; ++i) tmp[i] = T(0.5) * (ud[i] - ld[i]);
      element_op<T>(s, kOpBox, el, tmp.data(), e0.data(), nullptr);
      // G1's ramp: kap from the W-cube max / min of e0, gk = clip(Box_W ramp(kap), 0, 1)
      cube<T>(s, kCubeMax, e0.data(), kap.data(), W);
      cube<T>(s, kCubeMin, e0.data(), mn.data(), W);
      {
        const T k0 = static_cast<T>(c.ramp_k0);
        const double dd = c.ramp_k1 - c.ramp_k0;
        const T kden = static_cast<T>(dd >= 1e-12 ? dd : 1e-12);
        const T fl = static_cast<T>(tol >= c.tiny ? tol : c.tiny);
        for (int64_t i = 0; i < N; ++i) {
          const T hi = kap[i], lo = mn[i];
          const T sum = hi + lo;
          const T den = sum >= fl ? sum : fl;
          const T kp = (hi - lo) / den;
          kap[i] = clip<T>((kp - k0) / kden, T(0), T(1));
        }
      }
      cube<T>(s, kCubeBox, kap.data(), gk.data(), W);
      for (int64_t i = 0; i < N; ++i) gk[i] = clip<T>(gk[i], T(0), T(1));
      if (c.anorm_smooth <= 0) {
        e = e0;
      } else {
        for (int32_t it = 0; it < c.anorm_smooth; ++it) cube<T>(s, kCubeBox, it ? e.data() : e0.data(), e.data(), W);
      }
      std::vector<double> vals(e.begin(), e.end());
      const double e_med = select_median(vals);
      {
        double thr = c.anorm_floor * e_med;
        thr = thr >= tol ? thr : tol;
        thr = thr >= c.tiny ? thr : c.tiny;
        for (int64_t i = 0; i < N; ++i) e[i] = static_cast<double>(e[i]) >= thr ? e[i] : static_cast<T>(thr);
        if (tape_d != nullptr && j < kTapeDbls - 2) tape_d[2 + j - 1] = thr;
      }
      for (int64_t i = 0; i < N; ++i) q[i] = h[i] / e[i];
      envelope_mean<T>(s, el, q.data(), Mn.data(), u.data(), l.data());
      for (int64_t i = 0; i < N; ++i) {
        const T wv = gk[i];
        const T mA = clip<T>(e[i] * Mn[i], ld[i], ud[i]);
        Mn[i] = wv * mA + (T(1) - wv) * m[i];
      }
      uint64_t m_res = 0, m_h = 0;
      for (int64_t i = 0; i < N; ++i) {
        const T mm = a[i] * Mn[i];
        const T xv = h[i] - mm;
        hn[i] = xv;
        m_res = std::max(m_res, abs_bits<T>(r[i] - xv));
        m_h = std::max(m_h, abs_bits<T>(xv));
      }
      const bool accepted = !(bits_d(m_res) > r_inf) && !(bits_d(m_h) > r_inf);
      if (tape_i != nullptr) {
        tape_i[37] += 1;
        if (accepted && j < 31) tape_i[38] |= 1 << j;
      }
      if (accepted) {
        h.swap(hn);
        ++n_sifts;
      } else {
        sifting = false;
      }
    }
    // step 10: the commit (mode.cu:306-348,420-430; field_ops.cu:103-154)
    if (n_imfs + 1 >= cap) throw std::logic_error("rows full");
    T* row = sift_rows(n_imfs);
    for (int64_t i = 0; i < N; ++i) {
      row[i] = h[i];
      r[i] = r[i] - h[i];
    }
    out.window[n_imfs] = w;
    out.energy[n_imfs] = field_energy(s, h.data());
    out.spacing[n_imfs] = spacing;
    out.sifts[n_imfs] = n_sifts;
    ++n_imfs;
    amax = field_amax(s, r.data());
    tol = tie_tolerance<T>(tau, amax);
    w_prev = w;
    ++k;
    if (max_imf > 0 && n_imfs >= max_imf) status = kDone;
    if (!std::isfinite(amax)) fail(kErrNonFiniteResult);
    L.swap(Lp);
  }
  // finish (field_ops.cu:157-174; mode.cu:284-302)
  if (status != kError) {
    T* res = sift_rows(n_imfs);
    for (int64_t i = 0; i < N; ++i) res[i] = r[i];
    if (kexp != 0) {
      for (int64_t row = 0; row <= n_imfs; ++row) {
        T* p = sift_rows(row);
        for (int64_t i = 0; i < N; ++i) {
          T v;
          if constexpr (std::is_same_v<T, float>) v = static_cast<float>(std::ldexp(static_cast<double>(p[i]), kexp));
          else v = std::ldexp(p[i], kexp);
          p[i] = v;
          if (!std::isfinite(v)) err = kErrNonFiniteResult;
        }
      }
    }
  }
  if (status == kError || err != kErrNone) {
    out.error = err ? err : kErrInternal;
    out.n_imfs = 0;
  } else {
    out.n_imfs = n_imfs;
    for (int32_t j = 0; j < n_imfs; ++j) out.energy[j] = std::ldexp(out.energy[j], 2 * kexp);
  }
  return out;
}

}  // namespace orc

// ================================================================================================
// C ABI (ctypes from tests/ and bench.py). dtype: 0 float32, 1 float64. Strides are C order.
extern "C" {

int32_t oracle_rows_cap(int dim, const int64_t* n, int32_t max_imf) {
  int32_t n32[3] = {1, 1, 1};
  int64_t count = 1;
  for (int a = 0; a < dim; ++a) {
    n32[a] = static_cast<int32_t>(n[a]);
    count *= n[a];
  }
  const orc::Ctl c = orc::make_ctl(dim, n32, count, false);
  const int32_t bound = orc::mode_bound(c);
  const int32_t modes = max_imf > 0 ? std::min(bound, max_imf) : bound;
  return modes + 2;
}

// rows: batch * cap * count in dtype; hist_*: batch * (cap - 1). Returns 0, or -1 on an exception.
int oracle_decompose(int dtype, const void* x, int dim, int64_t batch, const int64_t* n, int32_t max_imf,
                     double tie, int32_t sifts, int32_t check_finite, int32_t threads, void* rows, int32_t* n_imfs,
                     int32_t* error, int32_t* hist_window, int32_t* hist_sifts, double* hist_spacing,
                     double* hist_energy) {
  try {
    const orc::Shape s = orc::make_shape(dim, n);
    orc::Ctl c = orc::make_ctl(dim, s.n, s.count, dtype == 0);
    c.sifts = sifts;
    const int32_t cap = oracle_rows_cap(dim, n, max_imf);
    const double tau = tie >= 0 ? tie : (dtype == 0 ? std::ldexp(1.0, -14) : std::ldexp(1.0, -30));
    orc::OrientPlan plan;
    if (dim >= 2) plan = orc::orient_plan(s);
    bool failed = false;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t b = 0; b < batch; ++b) {
      try {
        orc::ElemOut o;
        const size_t rb = static_cast<size_t>(b) * cap * s.count;
        if (dtype == 0) {
          o = orc::decompose_one<float>(s, c, static_cast<const float*>(x) + b * s.count, max_imf, tau,
                                        check_finite != 0, cap, static_cast<float*>(rows) + rb, &plan);
        } else {
          o = orc::decompose_one<double>(s, c, static_cast<const double*>(x) + b * s.count, max_imf, tau,
                                         check_finite != 0, cap, static_cast<double*>(rows) + rb, &plan);
        }
        n_imfs[b] = o.n_imfs;
        error[b] = o.error;
        for (int32_t j = 0; j < cap - 1; ++j) {
          const size_t h = static_cast<size_t>(b) * (cap - 1) + j;
          const bool keep = o.error == 0 && j < o.n_imfs;
          hist_window[h] = keep ? o.window[j] : 0;
          hist_sifts[h] = keep ? o.sifts[j] : 0;
          hist_spacing[h] = keep ? o.spacing[j] : 0.0;
          hist_energy[h] = keep ? o.energy[j] : 0.0;
        }
      } catch (...) {
#ifdef _OPENMP
#pragma omp critical
#endif
        failed = true;
      }
    }
    return failed ? -1 : 0;
  } catch (...) {
    return -1;
  }
}

// One element's decomposition with its tape (see TapeSink): rows cap * count; r, L: modes * count
// in dtype; lo: modes * count; ints: modes * 40; dbls: modes * 34. Returns 0, or -1 on an exception.
int oracle_decompose_tape(int dtype, const void* x, int dim, const int64_t* n, int32_t sifts, void* rows,
                          int32_t* n_imfs, int32_t* error, int32_t modes, void* r, void* L, uint8_t* lo,
                          int32_t* ints, double* dbls) {
  try {
    const orc::Shape s = orc::make_shape(dim, n);
    orc::Ctl c = orc::make_ctl(dim, s.n, s.count, dtype == 0);
    c.sifts = sifts;
    const int32_t cap = oracle_rows_cap(dim, n, -1);
    const double tau = dtype == 0 ? std::ldexp(1.0, -14) : std::ldexp(1.0, -30);
    orc::OrientPlan plan;
    if (dim >= 2) plan = orc::orient_plan(s);
    orc::TapeSink sink;
    sink.modes = modes;
    sink.r = r;
    sink.L = L;
    sink.lo = lo;
    sink.ints = ints;
    sink.dbls = dbls;
    orc::ElemOut o;
    if (dtype == 0)
      o = orc::decompose_one<float>(s, c, static_cast<const float*>(x), -1, tau, true, cap, static_cast<float*>(rows),
                                    &plan, &sink);
    else
      o = orc::decompose_one<double>(s, c, static_cast<const double*>(x), -1, tau, true, cap,
                                     static_cast<double*>(rows), &plan, &sink);
    *n_imfs = o.n_imfs;
    *error = o.error;
    return 0;
  } catch (...) {
    return -1;
  }
}

// The separable cube of one element: op 0 max, 1 min, 3 box.
void oracle_cube(int dtype, const void* in, int dim, const int64_t* n, int op, int32_t w, void* out) {
  const orc::Shape s = orc::make_shape(dim, n);
  if (dtype == 0) orc::cube<float>(s, op, static_cast<const float*>(in), static_cast<float*>(out), w);
  else orc::cube<double>(s, op, static_cast<const double*>(in), static_cast<double*>(out), w);
}

// One 1-D vHGW line (op 0 max, 1 min, 2 sum without the box scale).
void oracle_vhgw_line(int dtype, const void* in, int32_t n, int32_t w, int op, void* out) {
  if (dtype == 0) {
    std::vector<float> buf;
    orc::vhgw_line<float>(static_cast<const float*>(in), 1, n, w, static_cast<orc::Op>(op), static_cast<float*>(out), 1, buf);
  } else {
    std::vector<double> buf;
    orc::vhgw_line<double>(static_cast<const double*>(in), 1, n, w, static_cast<orc::Op>(op), static_cast<double*>(out), 1, buf);
  }
}

double oracle_median(double* v, int64_t n) {
  std::vector<double> a(v, v + n);
  return orc::select_median(a);
}
double oracle_quantile(double* v, int64_t n, double q) {
  std::vector<double> a(v, v + n);
  return orc::select_quantile(a, q);
}

double oracle_noise_sigma(int dtype, const void* r, int dim, const int64_t* n, int per_axis) {
  const orc::Shape s = orc::make_shape(dim, n);
  if (dtype == 0) return orc::noise_sigma<float>(s, static_cast<const float*>(r), per_axis != 0);
  return orc::noise_sigma<double>(s, static_cast<const double*>(r), per_axis != 0);
}

// Energy (canonical pairwise) and amax of one element.
void oracle_energy(int dtype, const void* f, int dim, const int64_t* n, double* energy, double* amax) {
  const orc::Shape s = orc::make_shape(dim, n);
  if (dtype == 0) {
    *energy = orc::field_energy<float>(s, static_cast<const float*>(f));
    *amax = orc::field_amax<float>(s, static_cast<const float*>(f));
  } else {
    *energy = orc::field_energy<double>(s, static_cast<const double*>(f));
    *amax = orc::field_amax<double>(s, static_cast<const double*>(f));
  }
}

void oracle_sig_flags(int dtype, const void* r, const void* tau, int dim, const int64_t* n, uint8_t* flags,
                      uint64_t* ext) {
  const orc::Shape s = orc::make_shape(dim, n);
  if (dtype == 0) orc::sig_flags<float>(s, static_cast<const float*>(r), static_cast<const float*>(tau), flags, ext);
  else orc::sig_flags<double>(s, static_cast<const double*>(r), static_cast<const double*>(tau), flags, ext);
}

// The threshold field; returns 1 when finite.
int oracle_tau(int dtype, const void* r, int dim, const int64_t* n, int32_t Wd, int32_t Wl, double sigma_n,
               double tol, void* tau) {
  const orc::Shape s = orc::make_shape(dim, n);
  orc::Ctl c = orc::make_ctl(dim, s.n, s.count, dtype == 0);
  if (dtype == 0) {
    std::vector<float> tmp(static_cast<size_t>(s.count));
    return orc::tau_field<float>(s, c, static_cast<const float*>(r), Wd, Wl, sigma_n, tol, nullptr,
                                 static_cast<float*>(tau), tmp.data());
  }
  std::vector<double> tmp(static_cast<size_t>(s.count));
  return orc::tau_field<double>(s, c, static_cast<const double*>(r), Wd, Wl, sigma_n, tol, nullptr,
                                static_cast<double*>(tau), tmp.data());
}

void oracle_raw_window(int dtype, const uint8_t* flags, int dim, const int64_t* n, int32_t win, void* om) {
  const orc::Shape s = orc::make_shape(dim, n);
  orc::Ctl c = orc::make_ctl(dim, s.n, s.count, dtype == 0);
  if (dtype == 0) {
    std::vector<float> tmp(static_cast<size_t>(s.count));
    orc::raw_window<float>(s, c, flags, win, static_cast<float*>(om), tmp.data());
  } else {
    std::vector<double> tmp(static_cast<size_t>(s.count));
    orc::raw_window<double>(s, c, flags, win, static_cast<double*>(om), tmp.data());
  }
}

void oracle_ccl(const uint8_t* mask, int dim, const int64_t* n, int32_t* labels) {
  const orc::Shape s = orc::make_shape(dim, n);
  std::vector<int32_t> lab;
  orc::ccl(s, mask, lab);
  std::copy(lab.begin(), lab.end(), labels);
}

// The orthonormal DCT-II of one element (double), and the inverse with the sin basis on the axes of odd.
void oracle_dct(const double* in, int dim, const int64_t* n, int inverse, int odd, double* out) {
  const orc::Shape s = orc::make_shape(dim, n);
  std::vector<double> f(in, in + s.count);
  for (int a = 0; a < dim; ++a) {
    const orc::AxisTables t = orc::axis_tables(s.n[a]);
    orc::axis_apply(s, a, !inverse ? t.fwd : (((odd >> a) & 1) ? t.isin : t.icos), f);
  }
  std::copy(f.begin(), f.end(), out);
}

// The orientation n (dim fields) and c of one element in the band of window w; returns 1 when finite.
int oracle_orientation(int dtype, const void* r, int dim, const int64_t* n, int32_t w, void* nvec, void* coh) {
  const orc::Shape s = orc::make_shape(dim, n);
  const orc::OrientPlan P = orc::orient_plan(s);
  if (dtype == 0) {
    std::vector<std::vector<float>> nv;
    std::vector<float> c;
    const bool ok = orc::orientation<float>(P, static_cast<const float*>(r), w, nv, c);
    for (int a = 0; a < dim; ++a) std::copy(nv[a].begin(), nv[a].end(), static_cast<float*>(nvec) + a * s.count);
    std::copy(c.begin(), c.end(), static_cast<float*>(coh));
    return ok;
  }
  std::vector<std::vector<double>> nv;
  std::vector<double> c;
  const bool ok = orc::orientation<double>(P, static_cast<const double*>(r), w, nv, c);
  for (int a = 0; a < dim; ++a) std::copy(nv[a].begin(), nv[a].end(), static_cast<double*>(nvec) + a * s.count);
  std::copy(c.begin(), c.end(), static_cast<double*>(coh));
  return ok;
}

// The monogenic signal of one element into out (3 planes for dim 1, 4 + dim otherwise, count samples
// each: amplitude, phase, frequency, the direction components, the coherence).
void oracle_monogenic(int dtype, const void* x, int dim, const int64_t* n, void* out) {
  const orc::Shape s = orc::make_shape(dim, n);
  const orc::OrientPlan P = orc::orient_plan(s);
  if (dtype == 0) orc::monogenic<float>(P, static_cast<const float*>(x), static_cast<float*>(out));
  else orc::monogenic<double>(P, static_cast<const double*>(x), static_cast<double*>(out));
}

// The Riesz components (R1: dim fields, R2: dim (dim + 1) / 2 fields) in double.
void oracle_riesz(const double* r, int dim, const int64_t* n, int32_t w, double* r1, double* r2) {
  const orc::Shape s = orc::make_shape(dim, n);
  const orc::OrientPlan P = orc::orient_plan(s);
  std::vector<std::vector<double>> a1, a2;
  orc::riesz<double>(P, r, w, a1, a2);
  for (size_t a = 0; a < a1.size(); ++a) std::copy(a1[a].begin(), a1[a].end(), r1 + a * s.count);
  for (size_t a = 0; a < a2.size(); ++a) std::copy(a2[a].begin(), a2[a].end(), r2 + a * s.count);
}

// n, c of symmetric tensors t (count x dim (dim + 1) / 2, pair order).
void oracle_eigen(int dim, const double* t, int64_t count, double* nvec, double* coh) {
  const int K = dim * (dim + 1) / 2;
  for (int64_t i = 0; i < count; ++i) orc::eigen(dim, t + i * K, nvec + i * dim, coh + i);
}

// The rung ladder; returns the rung count (-1: too many rungs) and writes the sides into out[32].
int oracle_rung_ladder(double d_min, double d_max, int32_t w, int32_t n_max, int32_t* out) {
  try {
    const std::vector<int32_t> v = orc::rung_ladder(d_min, d_max, w, n_max);
    std::copy(v.begin(), v.end(), out);
    return static_cast<int>(v.size());
  } catch (...) {
    return -1;
  }
}

// D-linear interpolation of f (double) at points y (count x dim) with periodic half-sample reflection.
void oracle_interp(const double* f, int dim, const int64_t* n, const double* y, int64_t count, double* out) {
  const orc::Shape s = orc::make_shape(dim, n);
  for (int64_t i = 0; i < count; ++i) out[i] = orc::interp<double>(s, f, y + i * dim);
}

}  // extern "C"
