// SPDX-License-Identifier: Apache-2.0
// The tape of a decomposition: the decisions of every mode that the adjoint holds frozen, kept when
// a call asks for it (eda_params.keep_tape) and read by the backward (adjoint.hip).
//
// The sift map of a mode takes the residual r at the mode's start to the mode h. Frozen are what
// the method decides from discrete data (flags, labels, selections, the guard's verdicts): the
// windows w (the band of the orientation) and W (G1's cubes), the scale field L (a function of the
// extrema flags and components only), the rung ladder with the rungs each element uses and those
// that carry segment points, the lower rung lo of every voxel, the tie tolerance, which sifts ran
// and which were accepted, and each sift's amplitude floor thr = max(0.1 median(e), tol, tiny).
// Everything else (the orientation, the element, the hat weights, every cube, gather and pointwise
// stage) is a function of r that the backward recomputes and differentiates.
#pragma once

#include "common.hpp"
#include "kernels.hpp"

#include <memory>
#include <vector>

namespace eda {

// Device memory owned by the tape (freed with it).
struct TapeBuf {
  void* p = nullptr;
  TapeBuf() = default;
  explicit TapeBuf(size_t bytes);
  TapeBuf(const TapeBuf&) = delete;
  TapeBuf& operator=(const TapeBuf&) = delete;
  TapeBuf(TapeBuf&& o) noexcept : p(o.p) { o.p = nullptr; }
  TapeBuf& operator=(TapeBuf&& o) noexcept {
    std::swap(p, o.p);
    return *this;
  }
  ~TapeBuf();
};

// One mode of a chunk: per element (index b < B of the chunk) and per sift j >= 1 (index j - 1).
struct TapeMode {
  std::vector<int32_t> active;  // the element runs (and commits) this mode
  std::vector<int32_t> w, W;
  std::vector<double> tol;
  std::vector<ElemDev> el;      // ladder, used and segment-point bits
  int32_t K_max = 0;
  std::vector<int32_t> rung_win;  // [K_max B]: rung j of element b (0: unused)
  std::vector<int32_t> rung_max;  // [K_max]
  std::vector<std::vector<int32_t>> ran, acc;  // [j - 1][B]
  std::vector<std::vector<double>> thr;        // [j - 1][B]
  TapeBuf r, L, lo;  // the residual and the scale field (T, B count), lo (uint8, B count)
};

struct ChunkTape {
  int64_t b0 = 0;
  Geom g{};
  std::vector<int32_t> kexp, n_imfs, error;  // [B]
  std::vector<TapeMode> modes;
};

struct Tape {
  int dtype = 0;
  int dct_mode = 0;  // the call's DCT contraction mode (set_dct_mode), reused by the backward
  MethodCtl c{};
  std::vector<ChunkTape> chunks;
};

// The backward of one chunk: grad_x (B, count) from grad_rows (B, P, count), both device arrays of
// the chunk's elements (orient_grad: differentiate the orientation; else hold n and c fixed).
template <class T>
void adjoint_chunk(hipStream_t s, int device, const Tape& tape, const ChunkTape& ct, const T* grad_rows, int32_t P,
                   T* grad_x, bool orient_grad);
// The frozen forward of one chunk at x (B, count): rows (B, P, count).
template <class T>
void replay_chunk(hipStream_t s, int device, const Tape& tape, const ChunkTape& ct, const T* x, T* rows, int32_t P);

}  // namespace eda
