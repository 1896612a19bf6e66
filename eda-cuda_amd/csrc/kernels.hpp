// SPDX-License-Identifier: Apache-2.0
// The internal interface between the host driver (driver.hip) and the kernel files. Every function
// enqueues work on the stream s and never synchronizes. Per-element arrays are device arrays [B];
// a window 0 (or on[b] == 0) skips element b, whose outputs are left untouched.
#pragma once

#include "common.hpp"

#include <functional>
#include <string>

#include <cstdint>

namespace eda {

// Per-call constants of the method (reference cpp/src/method/common.cpp:14-63).
struct MethodCtl {
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
MethodCtl method_ctl(const Geom& g, bool f32);

// Runs f, mapping its exceptions onto the C ABI's return codes and message (include/eda_hip.h).
int abi_guard(char* err, size_t errlen, const std::function<void()>& f);

// ---- the kernel timer (the benchmark's live roofline figures) -----------------------------------
// While a family is enabled, every launch of its kernels is bracketed by two HIP events on its
// stream and its algorithmic bytes are counted (each distinct input field read once, each output
// written once, over the elements that run; the count of those comes from the host mirror of the
// per-element arrays, with no device read-back). Families (bits of the mask): 1 the cube filter's
// passes (k_vhgw_rows, k_vhgw, k_vhgw_stream, k_box, k_cube2), 2 the DCT's axis passes (k_dct_fft, the dense contraction),
// 4 the rung gathers (k_gather_all), 8 the significant-extrema tiles (k_flags_tile).
enum TimerFamily : int { kTimeVhgw = 1, kTimeDct = 2, kTimeGather = 4, kTimeFlags = 8 };
enum TimerTag : int {
  kTagRows = 0,     // k_vhgw_rows
  kTagStrided,      // k_vhgw (the staged pass)
  kTagBoxShort,     // k_box, windows up to 31 samples
  kTagBoxLong,      // k_box, longer windows
  kTagDct,          // k_dct_fft and the dense contraction
  kTagGather,       // k_gather_all
  kTagFlags,        // k_flags_tile
  kTagStream,       // k_vhgw_stream (strided axes, windows up to 63 samples)
  kTagCube2,        // k_cube2 (both passes of a 2-D cube, windows up to 31 samples)
  kTagCount
};
// ---- the launch-shape autotuner (driver.hip) -----------------------------------------------------
// Per device and key (a kernel family and the shape facts its launch shape depends on), the first
// launch times every candidate shape twice on the launch's stream (HIP events; every candidate
// writes the same bits, so the output stands whichever ran last) and keeps the fastest; later
// launches take it. With EDA_AUTOTUNE=0 in the environment, candidate `fallback` always.
int autotune(const std::string& key, int ncand, int fallback, hipStream_t s, const std::function<void(int)>& launch);
// The choices so far, as JSON {key: {"choice": c, "us": [min time per candidate]}}.
std::string autotune_report();
void timer_set(int mask);
bool timer_on();                                                            // family 1
bool timer_gemm_on();                                                       // family 2
bool timer_family(int family);
void timer_note(hipStream_t s, hipEvent_t* start);                          // before a launch
void timer_done(hipStream_t s, hipEvent_t start, double bytes, int tag);    // after it
void timer_read(double* ms, int64_t* launches, double* bytes);              // synchronizes, sums, clears
// Host mirrors of device parameter arrays (registered by the driver's uploads, cleared with the
// upload arena): the host copy of [dev, dev + bytes), or nullptr when the range is not mirrored.
// The timer counts bytes from them, and the cube passes skip launches of window classes that no
// element of the launch falls in.
void mirror_note(const void* dev, const void* host, size_t bytes);
void mirror_clear();
const void* mirror_host(const void* dev, size_t bytes);
// the number of non-zero entries of win[0 .. n), or -1 when the array is not mirrored
int64_t mirror_active(const int32_t* win, int64_t n);

// ---- the cube filter (vhgw.hip) ------------------------------------------------------------------
enum class CubeOp : int8_t { kMax = 0, kMin = 1, kMaxMin = 2, kBox = 3 };
// win [B]: one odd side per element (0: skipped); max_w: the largest side of the launch (host).
template <class T>
void enqueue_cube(hipStream_t s, const Geom& g, CubeOp op, const T* in, T* out0, T* out1, const int32_t* win,
                  int32_t max_w);
// One max (min) pass along `axis` for `rungs` window arrays: rung j reads in + j in_stride and writes
// out + j out_stride (the pass the cube runs per axis).
template <class T>
void enqueue_axis_pass(hipStream_t s, const Geom& g, int axis, bool min, const T* in, T* out, const int32_t* win,
                       int32_t max_w, int32_t rungs, int64_t in_stride, int64_t out_stride);
// The cubes of `rungs` window arrays (win + j B) of one input in one launch per pass, rung j into
// out0 + j stride (out1 + j stride); each rung's outputs are bitwise those of enqueue_cube. With
// pairs (kMaxMin), the last pass may write the max and min of a voxel side by side instead:
// pairs + j 2 stride + 2 v (+ 1); returns whether it did (out0 / out1 then hold intermediates).
// scratch (max, min, box; rungs fields of `stride`): the first pass's output, so that no pass of a
// 2-D cube runs in place (the streaming passes need that); its contents are then undefined.
template <class T>
bool enqueue_cube_rungs(hipStream_t s, const Geom& g, CubeOp op, const T* in, T* out0, T* out1, const int32_t* win,
                        int32_t rungs, int64_t stride, int32_t max_w, T* pairs = nullptr, T* scratch = nullptr);

// ---- reductions (reduce.hip) ---------------------------------------------------------------------
// Per element: energy[b] = the canonical pairwise sum of squares (double; nullptr: none),
// amax_bits[b] = the largest bit pattern of |v| as a double, nonfinite[b] = the count of inf / NaN
// samples (nullptr: none). rows_ws: a device scratch of (B * rows) doubles + (B * rows) u64.
template <class T>
void enqueue_elem_stats(hipStream_t s, const Geom& g, const T* f, const int32_t* on, double* energy,
                        uint64_t* amax_bits, uint64_t* nonfinite, void* rows_ws);
size_t elem_stats_ws_bytes(const Geom& g);

// ---- exact selections (select.hip) ---------------------------------------------------------------
// The value set of element b: kField: field[i] where mask == nullptr or (mask[i] & bits) != 0;
// kDiff: forward differences along the axes of the bit mask `axes`, computed in T; optionally
// |value - center[b]| in double. Values compare as doubles.
enum class SelSrc : int8_t { kField = 0, kDiff = 1 };
template <class T>
struct Selection {
  SelSrc src;
  const T* field;
  const uint8_t* mask;
  uint8_t bits;
  uint8_t axes;
  const double* center;
};
// mode 0: median (numpy.median); 1: quantile q (numpy 'linear'). out[b]: NaN for an empty set;
// count[b] (nullptr: none): the size of the set.
// slots: the candidate slots per element (-1: select_cand_slots(g); 0: every pass reads every
// value); ws: select_ws_bytes(g, slots) bytes.
template <class T>
void enqueue_select(hipStream_t s, const Geom& g, const Selection<T>& sel, int mode, double q, double* out,
                    int64_t* count, const int32_t* on, void* ws, int64_t slots = -1);
int64_t select_cand_slots(const Geom& g);
size_t select_ws_bytes(const Geom& g, int64_t slots = -1);
// Noise levels: (1.4826 MAD) / sqrt 2 of the forward differences, pooled over the axes or the
// largest per axis. ws: select_ws_bytes + 3 [B] doubles.
template <class T>
void enqueue_noise_sigma(hipStream_t s, const Geom& g, const T* r, bool per_axis, double* out, const int32_t* on,
                         void* ws);
size_t noise_ws_bytes(const Geom& g);

// ---- significant extrema (flags.hip) -------------------------------------------------------------
// Clears flags of the running elements, then along every axis with n >= 3 sets bit a at the
// significant peaks of r and -r and adds their count into ext[3 b + a] (the caller zeroes ext).
// scratch: one field of T.
template <class T>
void enqueue_sig_flags(hipStream_t s, const Geom& g, const T* r, const T* tau, uint8_t* flags, uint64_t* ext,
                       const int32_t* on, T* scratch);

// ---- components (ccl.hip) ------------------------------------------------------------------------
// Face-connected components of {mask & bits}: labels[i] = the smallest in-element index of the
// component, -1 outside.
void enqueue_ccl(hipStream_t s, const Geom& g, const uint8_t* mask, uint8_t bits, int32_t* labels,
                 const int32_t* on);

// ---- orientation (orient.hip) --------------------------------------------------------------------
// The DCT contraction mode of the calling thread (0 exact float32; 1 split-BF16, three terms on
// the matrix cores, float32 calls only); returns the previous mode.
int set_dct_mode(int mode);
// The dense DCT matrices of one axis length, on the device (built once per (n, dtype)).
template <class T>
struct AxisMats {
  int32_t n = 0;
  T* fwd = nullptr;   // fwd[k * n + x] = c_k cos(pi k (2x+1) / (2n))
  T* icos = nullptr;  // icos[x * n + k] = c_k cos(...)
  T* isin = nullptr;  // isin[x * n + k] = c_k sin(...)
};
template <class T>
const AxisMats<T>& axis_mats(int device, int32_t n);
// n (D fields, component-major [a][B][count]) and coh of every element with on[b]; w [B]: the mode
// windows (the band's centre period). work: (3 + D (D + 1) / 2) fields of T; nonfinite[b] = 1
// where the tensor is not finite; dscratch: orient_scratch_doubles(B) doubles.
template <class T>
void enqueue_orientation(hipStream_t s, int device, const Geom& g, const T* r, const int32_t* w, const int32_t* on,
                         T* nvec, T* coh, T* work, int32_t* nonfinite, double* dscratch);
inline int64_t orient_scratch_doubles(int64_t batch) { return batch * 65; }
// The adjoint of the orientation: gr += (dn / dr)^T gn + (dc / dr)^T gc for the n and c that
// enqueue_orientation computes from r (D >= 2). work: D + D (D + 1) / 2 + 4 fields of T;
// dscratch: orient_scratch_doubles(B) doubles.
template <class T>
void enqueue_orientation_vjp(hipStream_t s, int device, const Geom& g, const T* r, const int32_t* w, const int32_t* on,
                             const T* gn, const T* gc, T* gr, T* work, double* dscratch);
// The monogenic signal of every element of f (the prescaled fields: the input times 2^-kexp[b])
// into the planes of out, the elements at b0 .. b0 + B - 1 of out_batch: amplitude (times
// 2^kexp[b]), phase, frequency [out_batch][count] each, then for D >= 2 the direction
// [out_batch][D][count] and the coherence [out_batch][count]. work: monogenic_fields(D) fields;
// dscratch: orient_scratch_doubles(B) doubles.
template <class T>
void enqueue_monogenic(hipStream_t s, int device, const Geom& g, const T* f, const int32_t* kexp, double tiny,
                       T* work, double* dscratch, T* out, int64_t out_batch, int64_t b0);
inline int monogenic_fields(int D) { return D >= 2 ? 3 + 3 * D + D * (D + 1) / 2 + 1 : 6; }
inline int monogenic_planes(int D) { return D == 1 ? 3 : 4 + D; }
// The orthonormal DCT-II of every element (inverse = 0) or its inverse with the sin basis on the
// axes of the bit mask odd (inverse = 1), through the same GEMMs; tmp: one field.
template <class T>
void enqueue_dct(hipStream_t s, int device, const Geom& g, const T* in, T* out, T* tmp, int inverse, int odd);
// n, c of the symmetric tensors in t (K = D (D + 1) / 2 fields, pair order) of the running elements.
template <class T>
void enqueue_eigen(hipStream_t s, const Geom& g, const T* t, const int32_t* on, T* nvec, T* coh, int32_t* nonfinite);

// ---- the element (element.hip) -------------------------------------------------------------------
// Per-element data of the structuring element (device, [B]).
struct ElemDev {
  int32_t n_rungs;
  uint32_t used, pts;
  int32_t rung_w[kMaxRungs];
};
// d, e, o (D >= 2) of every voxel and the per-element bits of min d / max d (u64 [2 B], the caller
// sets +inf / 0 first).
template <class T>
void enqueue_element_build(hipStream_t s, const Geom& g, const MethodCtl& c, const T* L, const T* nvec, const T* coh,
                           T* o, const int32_t* on, uint64_t* dminmax);
// lo, t of every voxel and the rung bits used / pts (u32 [2 B], atomicOr; the caller zeroes them).
template <class T>
void enqueue_element_weights(hipStream_t s, const Geom& g, const MethodCtl& c, const T* L, const T* coh,
                             const ElemDev* el, uint8_t* lo, T* t, const int32_t* on, uint32_t* bits);
// The operators over the rungs: op 0 Dil (max), 2 Dil + Ero (max into out0, min into out1), 3 Avg.
// rung_win [K_max * B]: rung j of element b at rung_win[j B + b] (0: unused); rung_max [K_max]:
// the largest side of rung j over the batch (host). The cubes of every rung run in one launch per
// pass into rungs0 (rungs1: the min cubes of op 2), K_max fields each; then one gather over all
// rungs (op 2 reads the max and min side by side from pairs, 2 K_max fields, when the cube wrote
// them there). on [B]: the elements that run (the others' outputs are left untouched).
// Fields the gather of an element operator writes beside its outputs (each optional): of the max
// u and min l (op 2) mid = 0.5 (u + l), hdiff = 0.5 (u - l), pdiff = max(u - l, 0); of the box m
// (op 3) diff = minuend - m.
template <class T>
struct GatherExtra {
  T* mid = nullptr;
  T* hdiff = nullptr;
  T* pdiff = nullptr;
  const T* minuend = nullptr;
  T* diff = nullptr;
};

template <class T>
void enqueue_element_op(hipStream_t s, const Geom& g, int op, const ElemDev* el, const uint8_t* lo, const T* t,
                        const T* o, const int32_t* rung_win, const int32_t* rung_max, int K_max, const T* in, T* out0,
                        T* out1, T* rungs0, T* rungs1, T* pairs, const int32_t* on, GatherExtra<T> extra = {});

}  // namespace eda
