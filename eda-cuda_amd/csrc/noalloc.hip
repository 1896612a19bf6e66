// SPDX-License-Identifier: Apache-2.0
// The allocation-free order statistic (eda_prim_select_noalloc, include/eda_hip.h).
//
// Callers that cannot allocate device memory (code under HIP graph capture, serving loops that
// forbid hipMalloc) still need the selection's workspace: its per-element search state, the two
// 2048-bin histograms and the candidate lists. This entry point takes them from `scratch`, a
// file-scope device array of kScratchBytes that every GPU the library runs on holds until the
// process ends. The array's name is a host-side stand-in: host code uses only the device address
// that hipGetSymbolAddress gives on each device (looked up once per device, with that device
// current, at the entry point's first call there). Nothing here calls hipMalloc, hipFree, a
// library routine that allocates, or the per-thread context of the other entry points (whose set-up
// allocates); the kernels run on the caller's stream and the call does not wait for them.
//
// Every call on one GPU shares that GPU's array, and the array keeps the last call's contents (the
// selection clears what must start at zero with a kernel, not hipMemsetAsync): calls that overlap
// on one GPU (two streams, or two threads) race and give undefined results.
#include "common.hpp"
#include "kernels.hpp"
#include "../../include/eda_hip.h"

#include <mutex>

namespace eda {
namespace {

constexpr size_t kScratchBytes = size_t{32} << 20;  // 32 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up on its first use there).
void* scratch_address() {
  int dev = 0;
  EDA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::invalid_argument("eda: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    EDA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (p == nullptr) throw std::runtime_error("eda: no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

template <class T>
void select_noalloc(const Geom& g, const Selection<T>& sel, int32_t mode, double q, double* out, int64_t* count,
                    hipStream_t s) {
  // the candidate lists take what the state and histograms leave, up to their usual size
  const size_t fixed = select_ws_bytes(g, 0);
  if (fixed > kScratchBytes) {
    throw std::invalid_argument("eda: select_noalloc: " + std::to_string(g.batch) + " elements need " +
                                std::to_string(fixed >> 10) + " KiB of scratch, more than its " +
                                std::to_string(kScratchBytes >> 10) + " KiB");
  }
  int64_t slots = select_cand_slots(g);
  const int64_t room = static_cast<int64_t>((kScratchBytes - fixed) / sizeof(uint64_t)) / g.batch - 32;
  if (slots > room) slots = room;
  if (slots < 64) slots = 0;
  if (select_ws_bytes(g, slots) > kScratchBytes) throw std::logic_error("eda: select_noalloc: scratch layout");
  enqueue_select<T>(s, g, sel, mode, q, out, count, nullptr, scratch_address(), slots);
}

}  // namespace
}  // namespace eda

using namespace eda;

extern "C" int eda_prim_select_noalloc(int32_t dtype, int32_t dim, int64_t batch, const int64_t* n,
                                       const void* field, const uint8_t* mask, int32_t bits, int32_t mode, double q,
                                       double* out, int64_t* count, uint64_t stream, char* err, size_t errlen) {
  return abi_guard(err, errlen, [&] {
    if (dtype != EDA_F32 && dtype != EDA_F64) throw std::invalid_argument("eda: dtype must be 0 or 1");
    if (dim < 1 || dim > 3 || batch < 1 || n == nullptr) throw std::invalid_argument("eda: prim: bad geometry");
    if (field == nullptr || out == nullptr) throw std::invalid_argument("eda: select_noalloc: null pointer");
    if (mode != 0 && !(q >= 0.0 && q <= 1.0)) throw std::invalid_argument("eda: quantile: q must lie in [0, 1]");
    const Geom g = make_geom(dim, batch, n);
    const hipStream_t s = reinterpret_cast<hipStream_t>(static_cast<uintptr_t>(stream));
    if (dtype == EDA_F32) {
      Selection<float> sel{};
      sel.src = SelSrc::kField;
      sel.field = static_cast<const float*>(field);
      sel.mask = mask;
      sel.bits = static_cast<uint8_t>(bits);
      select_noalloc<float>(g, sel, mode, q, out, count, s);
    } else {
      Selection<double> sel{};
      sel.src = SelSrc::kField;
      sel.field = static_cast<const double*>(field);
      sel.mask = mask;
      sel.bits = static_cast<uint8_t>(bits);
      select_noalloc<double>(g, sel, mode, q, out, count, s);
    }
  });
}
