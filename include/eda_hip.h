/* SPDX-License-Identifier: Apache-2.0
 *
 * The C ABI of the MI355X (gfx950) decomposition library libeda_hip.so: plain pointers and sizes,
 * no C++ or framework types. It is the drop-in boundary of `eda.decompose`: the reference binds the
 * same calls through its nanobind module `eda._core`; the Python package eda (ctypes) and any other
 * FFI bind these entry points instead.
 *
 * Conventions: every entry point returns EDA_OK or one of the error codes below and writes a message
 * starting with "eda:" into err (err may be NULL). Device pointers are HIP device pointers of the
 * device named in the call; host pointers are ordinary memory. Calls from several threads are
 * supported: each (thread, device) pair has its own stream and workspace.
 */
#ifndef EDA_HIP_H
#define EDA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes (reference cpp/python/module.cpp:28-51 maps its C++ exceptions onto the same
 * Python classes): */
#define EDA_OK 0
#define EDA_EINVAL 1    /* an invalid argument: Python ValueError */
#define EDA_ENOMEM 2    /* host or device memory exhausted: Python MemoryError */
#define EDA_EHIP 3      /* a HIP runtime call or a kernel launch failed: Python eda.CudaError */
#define EDA_EINTERNAL 4 /* an internal invariant failed: Python RuntimeError */

/* dtype codes */
#define EDA_F32 0
#define EDA_F64 1

/* Element types of device input (eda_array_desc.source; reference cpp/src/field/convert.hpp:17-22).
 * 0: the compute dtype itself. Each element is cast to the compute dtype with round to nearest even
 * (bool to 0 or 1; float16, bfloat16 and float32 to float64 exactly). */
#define EDA_SRC_SAME 0
#define EDA_SRC_BOOL 1
#define EDA_SRC_I8 2
#define EDA_SRC_I16 3
#define EDA_SRC_I32 4
#define EDA_SRC_I64 5
#define EDA_SRC_U8 6
#define EDA_SRC_U16 7
#define EDA_SRC_U32 8
#define EDA_SRC_U64 9
#define EDA_SRC_F16 10
#define EDA_SRC_BF16 11
#define EDA_SRC_F32 12
#define EDA_SRC_F64 13

/* The input of a call: `batch` elements of `dim` (1..3) spatial axes. Host input (device = -1) is
 * C-contiguous (batch, n[0], .., n[dim-1]) in the compute dtype. Device input is read in place,
 * element (b, i) at data[offset(b) + sum_a i_a strides[a]] (element strides of the source type, any
 * sign), with offset(b) = b batch_stride, or, when batch_axes > 1, sum_k j_k batch_strides[k] for
 * j the C-order multi-index of b over batch_shape[0 .. batch_axes) (reference
 * cpp/src/decompose.cpp:25-82, the strided gather of cpp/src/field/convert.hpp:33-49). */
typedef struct eda_array_desc {
  const void* data;
  int32_t device; /* -1: host memory; else the HIP device that holds data */
  int32_t dtype;  /* EDA_F32 or EDA_F64: the compute dtype */
  int32_t dim;
  int32_t source; /* device input: its element type, EDA_SRC_* (EDA_SRC_SAME: the compute dtype) */
  int64_t batch;
  int64_t n[3];
  int64_t batch_stride; /* in elements (device input) */
  int64_t strides[3];   /* in elements (device input) */
  int32_t batch_axes;   /* device input: > 1 to use batch_shape / batch_strides (at most 64 axes) */
  int32_t reserved;
  const int64_t* batch_shape;   /* host arrays of batch_axes entries (batch = their product) */
  const int64_t* batch_strides; /* in elements of the source type */
} eda_array_desc;

/* The parameters of a call (reference cpp/src/decompose.hpp:69-76, src/eda/_api.py:184-353). */
typedef struct eda_params {
  int32_t max_imf;      /* <= 0: no limit */
  int32_t sifts;        /* >= 1 */
  int32_t check_finite; /* 1: an element with inf / NaN input fails with error 1 */
  int32_t orientation;  /* 1: keep the orientation (n, c) of every mode (dim >= 2) for eda_collect_orientation */
  double tie_tolerance; /* < 0: the dtype's default (2^-14 float32, 2^-30 float64) */
  int64_t max_chunk;    /* > 0: at most this many elements per chunk; 0: by free memory */
  uint64_t stream;      /* a hipStream_t (0: the null stream) that device input is read after */
  int32_t keep_tape;    /* 1: keep every mode's frozen decisions in the result, for eda_backward / eda_replay */
  int32_t dct_mode;     /* float32 only: 0 the exact DCT / Riesz contractions (an LDS FFT for power-of-two
                           axes, float32 MFMA GEMMs otherwise); 1 every axis as a split-BF16 GEMM on the
                           matrix cores (3 products per term, float32 accumulation) */
} eda_params;

/* The result of eda_decompose. The rows stay in a library-owned device buffer until eda_collect
 * copies them out; eda_release frees everything. Info arrays are host arrays owned by the result:
 * n_imfs, error [batch]; windows, sifts, spacing, energy [batch * K] (K = max n_imfs; entries
 * beyond an element's n_imfs and all entries of a failed element are 0). */
typedef struct eda_result {
  void* handle;
  int64_t batch;
  int64_t count;  /* samples per element */
  int32_t rows;   /* P = max n_imfs + 1 over the elements without an error (>= 1 if any failed) */
  int32_t K;
  int32_t dtype;
  int32_t device;
  int64_t chunk;  /* elements per chunk */
  int32_t* n_imfs;
  int32_t* error; /* 0, 1 (non-finite input), 2 (non-finite intermediate or row), 255 (internal) */
  int32_t* windows;
  int32_t* sifts;
  double* spacing;
  double* energy;
} eda_result;

/* Decompose every element of x: modes (IMFs, fine to coarse) and the residue. Replaces the
 * reference's `eda._core.decompose_host` (cpp/python/decompose.cpp:325-384) and
 * `eda._core.decompose_device` (cpp/python/decompose.cpp:452-509), which both run
 * eda::decompose (cpp/src/decompose.hpp:106). */
int eda_decompose(const eda_array_desc* x, const eda_params* p, eda_result* out, char* err, size_t errlen);

/* Copies rows [0, P) of every element into dst, C-contiguous (batch, P, *spatial) in the compute
 * dtype: rows beyond an element's n_imfs are 0, the rows of a failed element NaN. dst_device -1:
 * host memory (synchronous); else device memory of that device, complete on return. Replaces
 * eda::Run::collect (cpp/src/decompose.hpp:100-101). */
int eda_collect(const eda_result* r, void* dst, int32_t dst_device, char* err, size_t errlen);

/* With eda_params.orientation: copies the direction n (batch, K, dim, *spatial) into dir and the
 * coherence c (batch, K, *spatial) into coh, C-contiguous in the compute dtype: row k of element b
 * is n and c as the method computed them at the start of its mode k, n with the canonical sign
 * (its first component of magnitude above the dtype's default tie tolerance is positive); rows
 * beyond an element's n_imfs and every row of a failed element are 0. dst_device as in
 * eda_collect. Nothing is copied when batch or K is 0; a call that kept no orientation (orientation
 * 0, or dim 1) fails with EDA_EINTERNAL. Replaces eda::Run::collect_orientation
 * (cpp/src/decompose.hpp:100-103, cpp/src/method/driver.cpp:177-196). */
int eda_collect_orientation(const eda_result* r, void* dir, void* coh, int32_t dst_device, char* err, size_t errlen);

/* The monogenic signal of every element of x (reference `eda._core.monogenic_host` /
 * `monogenic_device`, cpp/python/monogenic.cpp; cpp/src/method/monogenic.cpp:61-158,
 * orientation.cu:420-493,789-831). out: the planes, C-contiguous in the compute dtype, one after the
 * other: amplitude, phase, frequency (batch, *spatial) each, then for dim >= 2 the direction
 * (batch, dim, *spatial) and the coherence (batch, *spatial): 3 planes for dim 1, 4 + dim
 * otherwise. out_device -1: host memory; else device memory of that device. nonfinite (host
 * [batch], may be NULL): 1 where check_finite and the element holds inf or NaN (its outputs are
 * still computed). chunk (may be NULL): the elements per chunk. max_chunk > 0 bounds it; stream: a
 * hipStream_t the call waits for before it reads device input (0: the null stream). Complete on
 * return. */
int eda_monogenic(const eda_array_desc* x, int32_t check_finite, int64_t max_chunk, uint64_t stream, void* out,
                  int32_t out_device, int32_t* nonfinite, int64_t* chunk, char* err, size_t errlen);

/* The adjoint of a decomposition whose call kept its tape (eda_params.keep_tape = 1): with the
 * decisions of every mode frozen (windows, rung ladders and the rungs used, the scale field, the
 * lower rung of every voxel, the sifts run and accepted, the amplitude floors), grad_x (batch,
 * *spatial) = J^T grad_rows, J the Jacobian of the rows (batch, P, *spatial) in the input, both
 * C-contiguous device arrays in the compute dtype on the result's device; complete on return. flags
 * bit 0: hold the orientation (direction and coherence) fixed instead of differentiating it. Failed
 * elements get 0. Axes longer than 2048 samples are refused (EDA_EINVAL). No counterpart in the
 * reference (it has no adjoint); the autograd form of eda.decompose sits on it. */
int eda_backward(const eda_result* r, const void* grad_rows, void* grad_x, int32_t flags, char* err, size_t errlen);

/* The frozen forward of a result with a tape at a new input x (batch, *spatial): rows (batch, P,
 * *spatial), device arrays as in eda_backward; at the call's own input it gives the call's rows.
 * For checking the adjoint against finite differences. No counterpart in the reference. */
int eda_replay(const eda_result* r, const void* x, void* rows, char* err, size_t errlen);

/* The frozen decisions of one mode of one element of a result with a tape: ints[40] = [w, W, K,
 * used bits, segment-point bits, rung_w[0..31], sifts run after the first, accepted bits (bit j:
 * sift j), the prescale exponent], dbls[34] = [tol, NaN, thr_j at index 1 + j]; and, into host
 * buffers of count samples where not null, the residual at the mode's start, the scale field L
 * (the result's dtype) and the lower rung of every voxel (uint8). For checking the decisions
 * against the oracle's; no counterpart in the reference. */
int eda_tape_mode(const eda_result* r, int64_t element, int32_t mode, int32_t* ints, double* dbls, void* residual,
                  void* L, uint8_t* lo, char* err, size_t errlen);

/* Frees the rows and the info arrays of a result (idempotent). */
void eda_release(eda_result* r);

/* Replaces eda._core.build_info / device_count / device_info / pointer_device
 * (cpp/python/module.cpp:64-142): */
int eda_build_info(char* buf, size_t len);                   /* a JSON object */
int eda_device_count(void);                                  /* 0 without a driver or device */
int eda_device_info(int32_t device, char* buf, size_t len);  /* a JSON object */
int eda_pointer_device(const void* ptr, int32_t* device);    /* -1 for host memory */

/* Kernel timing, for the live roofline figures of a benchmark. `on` is a mask of kernel families:
 * 1 the cube filter's pass kernels (k_vhgw_rows, k_vhgw, k_box), 2 the DCT's axis passes, 4 the
 * rung gathers, 8 the significant-extrema tiles; 0 turns timing off. While a family is on, every
 * launch of its kernels is bracketed by HIP events on the stream it runs on and its algorithmic
 * bytes are counted (each distinct input field read once, each output field written once, over the
 * running elements; a DCT axis pass one read and one write of the field). eda_timer_read waits for
 * the recorded launches, returns their total kernel time in ms, the launch count and the bytes, and
 * clears them (their per-kernel sums are kept for eda_timer_report). eda_timer_report writes the
 * per-kernel sums since the last report as a JSON object {kernel: {"ms", "launches", "bytes"}}
 * into buf and clears them. No counterpart in the reference (it has no timing in code). */
int eda_timer_enable(int32_t on);
int eda_timer_read(double* ms, int64_t* launches, double* bytes, char* err, size_t errlen);
int eda_timer_report(char* buf, size_t len, char* err, size_t errlen);
/* The launch shapes the autotuner chose so far (per device and kernel shape key: the candidate
 * index and each candidate's best time in us), as a JSON object; the first launch of each key
 * times its candidates (EDA_AUTOTUNE=0 in the environment turns that off). No counterpart in the
 * reference (it sizes its launches by fixed rules). */
int eda_autotune_report(char* buf, size_t len, char* err, size_t errlen);

/* ---- Primitive entry points (device pointers; each call synchronizes before it returns). They
 * expose the method's stages one by one, so that each can be checked against its closed form and
 * against the CPU oracle; the decomposition itself does not call them. Fields are C-contiguous
 * (batch, n[0], .., n[dim-1]). ---- */

/* The separable cube filter of every element with side win[b] (host array; 0 skips the element):
 * op 0 max, 1 min, 2 max into out0 and min into out1, 3 box mean (reference
 * cpp/src/method/common.cpp:65-136, enqueue_cube). */
int eda_prim_cube(int32_t dtype, int32_t dim, int64_t batch, const int64_t* n, int32_t op, const int32_t* win,
                  const void* in, void* out0, void* out1, char* err, size_t errlen);
/* Per element: the canonical energy, max |v| and the count of non-finite samples (host outputs;
 * reference cpp/src/field/field_ops.cu:52-83). */
int eda_prim_stats(int32_t dtype, int32_t dim, int64_t batch, const int64_t* n, const void* f, double* energy,
                   double* amax, uint64_t* nonfinite, char* err, size_t errlen);
/* The median (mode 0) or the linear quantile q (mode 1) of every element's values where
 * mask & bits != 0 (mask NULL: all), as numpy; out, count: host [batch] (reference
 * cpp/src/method/stats.cu:261-295). */
int eda_prim_select(int32_t dtype, int32_t dim, int64_t batch, const int64_t* n, const void* field,
                    const uint8_t* mask, int32_t bits, int32_t mode, double q, double* out, int64_t* count, char* err,
                    size_t errlen);
/* The same order statistics as eda_prim_select without allocating anything: no counterpart in
 * the reference; for callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops that forbid hipMalloc). field, mask (may be NULL), out [batch] and count [batch]
 * (may be NULL) are device pointers of the current device. The workspace is a static 32 MiB device
 * array per GPU (held until the process ends); a call whose batch needs more for its fixed part
 * (about 33 KiB per element: at most about 990 elements) fails with EDA_EINVAL. The kernels run on
 * `stream` (a hipStream_t; 0: the null stream) and the call returns without waiting for them.
 * Calls that overlap on one GPU share the array and race: the caller orders them. */
int eda_prim_select_noalloc(int32_t dtype, int32_t dim, int64_t batch, const int64_t* n, const void* field,
                            const uint8_t* mask, int32_t bits, int32_t mode, double q, double* out, int64_t* count,
                            uint64_t stream, char* err, size_t errlen);
/* The noise level (1.4826 MAD / sqrt 2 of the forward differences), pooled or the largest per axis
 * (reference cpp/src/method/stats.cu:299-334); out: host [batch]. */
int eda_prim_noise(int32_t dtype, int32_t dim, int64_t batch, const int64_t* n, const void* r, int32_t per_axis,
                   double* out, char* err, size_t errlen);
/* The threshold field with boxes Wd, Wl and per-element sigma_n, tol (host arrays); nonfinite:
 * host [batch] (reference cpp/src/method/stats_fields.cu:375-389). */
int eda_prim_tau(int32_t dtype, int32_t dim, int64_t batch, const int64_t* n, const void* r, const int32_t* Wd,
                 const int32_t* Wl, const double* sigma_n, const double* tol, void* tau, int32_t* nonfinite, char* err,
                 size_t errlen);
/* Significant extrema flags (bit a per axis) and their counts ext: host [batch * 3] (reference
 * cpp/src/method/stats_fields.cu:392-422). */
int eda_prim_flags(int32_t dtype, int32_t dim, int64_t batch, const int64_t* n, const void* r, const void* tau,
                   uint8_t* flags, uint64_t* ext, char* err, size_t errlen);
/* The raw local window of the flags with box win[b] (host) (reference stats_fields.cu:425-440). */
int eda_prim_raw_window(int32_t dtype, int32_t dim, int64_t batch, const int64_t* n, const uint8_t* flags,
                        const int32_t* win, void* om, char* err, size_t errlen);
/* Face-connected components of mask & bits: labels = the smallest in-element index of the
 * component, -1 outside (reference cpp/src/method/stats_scale.cu:208-220). */
int eda_prim_ccl(int32_t dim, int64_t batch, const int64_t* n, const uint8_t* mask, int32_t bits, int32_t* labels,
                 char* err, size_t errlen);
/* The orthonormal N-D DCT-II (inverse 0) or its inverse with the sin basis on the axes of the bit
 * mask odd (inverse 1) of every element (the reference's DctPlans, cpp/src/field/dct.hpp:35-85). */
int eda_prim_dct(int32_t dtype, int32_t dim, int64_t batch, const int64_t* n, const void* in, void* out,
                 int32_t inverse, int32_t odd, char* err, size_t errlen);
/* The orientation n (dim fields, [a][batch][count]) and the coherence of every element in the band
 * of window w[b] (host) (reference cpp/src/method/orientation.cu:771-778). */
int eda_prim_orientation(int32_t dtype, int32_t dim, int64_t batch, const int64_t* n, const void* r, const int32_t* w,
                         void* nvec, void* coh, char* err, size_t errlen);

#ifdef __cplusplus
}
#endif

#endif /* EDA_HIP_H */
