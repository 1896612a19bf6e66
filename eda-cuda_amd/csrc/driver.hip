// SPDX-License-Identifier: Apache-2.0
// The decomposition driver and the C ABI (include/eda_hip.h).
//
// A call runs its elements in chunks; a chunk runs the method's modes in lockstep on one HIP stream
// (reference cpp/src/method/driver.cpp:393-700). Field work is device kernels: the cube filter,
// selections, extrema, components, the orientation and the element's operators are in their own
// files; the pointwise stages of the threshold, the scale field, the continuation weight, G1, the
// sifts and the commit are device lambdas here (each() of common.hpp). The per-element control of a
// mode (the window rule, G2's rung machine, the growth floor, the ladder, the sift guard's verdict,
// the commit's bookkeeping) runs on the host between synchronizations, in the reference's order,
// from per-element statistics the kernels reduce on the device.
//
// Per-element arrays that kernels read (windows, thresholds, run flags) are uploaded through a
// pinned staging area into a device parameter arena, reset after every synchronization.
#include "common.hpp"
#include "kernels.hpp"
#include "tape.hpp"
#include "../../include/eda_hip.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <atomic>
#include <vector>

namespace eda {

void throw_hip(hipError_t e, const char* expr, const char* file, int line) {
  char buf[512];
  std::snprintf(buf, sizeof buf, "eda: HIP error %d (%s) in %s at %s:%d", static_cast<int>(e), hipGetErrorName(e), expr,
                file, line);
  if (e == hipErrorOutOfMemory) throw OutOfDeviceMemory(buf);
  throw HipError(e, buf);
}

// ---- the pass timer and the host mirrors of window arrays ----------------------------------------
namespace {
struct TimerState {
  std::mutex mu;
  std::atomic<int> on{0};  // a mask of TimerFamily bits
  std::vector<hipEvent_t> pool;
  struct Span {
    hipEvent_t a, b;
    double bytes;
    int tag;
  };
  std::vector<Span> spans;
  struct Total {
    double ms = 0.0, bytes = 0.0;
    int64_t launches = 0;
  };
  Total per_tag[kTagCount];  // summed by timer_read and kept until eda_timer_report
  hipEvent_t take() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    EDA_HIP(hipEventCreate(&e));
    return e;
  }
};
TimerState& timer_state() {
  static TimerState t;
  return t;
}
struct Mirror {
  const char* dev;
  const char* host;
  size_t bytes;
};
thread_local std::vector<Mirror> g_mirrors;
}  // namespace

void timer_set(int mask) { timer_state().on = mask; }
bool timer_family(int family) { return (timer_state().on.load(std::memory_order_relaxed) & family) != 0; }
bool timer_on() { return timer_family(kTimeVhgw); }
bool timer_gemm_on() { return timer_family(kTimeDct); }
void timer_note(hipStream_t s, hipEvent_t* start) {
  flush_uploads();  // a staged upload is not part of the timed launch
  TimerState& t = timer_state();
  std::lock_guard<std::mutex> lock(t.mu);
  *start = t.take();
  EDA_HIP(hipEventRecord(*start, s));
}
void timer_done(hipStream_t s, hipEvent_t start, double bytes, int tag) {
  TimerState& t = timer_state();
  std::lock_guard<std::mutex> lock(t.mu);
  hipEvent_t stop = t.take();
  EDA_HIP(hipEventRecord(stop, s));
  t.spans.push_back({start, stop, bytes, tag});
}
void timer_read(double* ms, int64_t* launches, double* bytes) {
  TimerState& t = timer_state();
  std::lock_guard<std::mutex> lock(t.mu);
  double sum = 0.0, by = 0.0;
  for (const auto& sp : t.spans) {
    EDA_HIP(hipEventSynchronize(sp.b));
    float e = 0.f;
    EDA_HIP(hipEventElapsedTime(&e, sp.a, sp.b));
    sum += e;
    by += sp.bytes;
    if (sp.tag >= 0 && sp.tag < kTagCount) {
      TimerState::Total& tt = t.per_tag[sp.tag];
      tt.ms += e;
      tt.bytes += sp.bytes;
      tt.launches += 1;
    }
    t.pool.push_back(sp.a);
    t.pool.push_back(sp.b);
  }
  *ms = sum;
  *launches = static_cast<int64_t>(t.spans.size());
  *bytes = by;
  t.spans.clear();
}
// ---- the autotuner ----
namespace {
struct TuneEntry {
  int choice = 0;
  std::vector<double> us;
};
struct TuneState {
  std::mutex mu;
  std::map<std::pair<int, std::string>, TuneEntry> table;
};
TuneState& tune_state() {
  static TuneState* t = new TuneState();  // outlives the static destructors of the process's end
  return *t;
}
bool autotune_on() {
  static const bool on = [] {
    const char* v = std::getenv("EDA_AUTOTUNE");
    return v == nullptr || std::strcmp(v, "0") != 0;
  }();
  return on;
}
}  // namespace

int autotune(const std::string& key, int ncand, int fallback, hipStream_t s, const std::function<void(int)>& launch) {
  if (fallback < 0 || fallback >= std::max(ncand, 1)) throw std::logic_error("eda: autotuner fallback outside its list");
  if (ncand <= 1 || !autotune_on()) {
    launch(fallback);
    return fallback;
  }
  int dev = 0;
  EDA_HIP(hipGetDevice(&dev));
  TuneState& t = tune_state();
  // the candidate count is part of the key: one shape key can have lists of different lengths
  // (the window bounds the piece counts), and a choice indexes its own list only
  const auto k = std::make_pair(dev, key + ":of" + std::to_string(ncand));
  {
    std::lock_guard<std::mutex> lock(t.mu);
    auto it = t.table.find(k);
    if (it != t.table.end()) {
      const int c = it->second.choice;
      if (c < 0 || c >= ncand) throw std::logic_error("eda: autotuner choice outside its candidate list");
      launch(c);
      return c;
    }
  }
  // two rounds over the candidates, each launch between two events; the minimum per candidate
  constexpr int kRounds = 2;
  std::vector<hipEvent_t> ev(static_cast<size_t>(2 * kRounds * ncand), nullptr);
  for (auto& e : ev) EDA_HIP(hipEventCreate(&e));
  for (int r = 0; r < kRounds; ++r)
    for (int c = 0; c < ncand; ++c) {
      flush_uploads();
      EDA_HIP(hipEventRecord(ev[2 * (r * ncand + c)], s));
      launch(c);
      EDA_HIP(hipEventRecord(ev[2 * (r * ncand + c) + 1], s));
    }
  EDA_HIP(hipEventSynchronize(ev.back()));
  TuneEntry entry;
  entry.us.assign(static_cast<size_t>(ncand), 1e30);
  for (int r = 0; r < kRounds; ++r)
    for (int c = 0; c < ncand; ++c) {
      float ms = 0.0f;
      EDA_HIP(hipEventElapsedTime(&ms, ev[2 * (r * ncand + c)], ev[2 * (r * ncand + c) + 1]));
      entry.us[c] = std::min(entry.us[c], 1e3 * static_cast<double>(ms));
    }
  for (auto& e : ev) (void)hipEventDestroy(e);
  entry.choice = static_cast<int>(std::min_element(entry.us.begin(), entry.us.end()) - entry.us.begin());
  const int c = entry.choice;
  std::lock_guard<std::mutex> lock(t.mu);
  t.table.emplace(k, std::move(entry));
  return c;
}

std::string autotune_report() {
  TuneState& t = tune_state();
  std::lock_guard<std::mutex> lock(t.mu);
  std::string js = "{";
  bool first = true;
  for (const auto& kv : t.table) {
    char head[64];
    std::snprintf(head, sizeof head, "%s\"%d:", first ? "" : ", ", kv.first.first);
    js += head + kv.first.second + "\": {\"choice\": " + std::to_string(kv.second.choice) + ", \"us\": [";
    for (size_t i = 0; i < kv.second.us.size(); ++i) {
      char num[32];
      std::snprintf(num, sizeof num, "%s%.2f", i ? ", " : "", kv.second.us[i]);
      js += num;
    }
    js += "]}";
    first = false;
  }
  return js + "}";
}

// the per-kernel totals summed by the timer_read calls since the last report, as JSON; clears them
std::string timer_report() {
  static const char* const names[kTagCount] = {"k_vhgw_rows", "k_vhgw", "k_box_short", "k_box_long",
                                                "k_dct_fft", "k_gather_all", "k_flags_tile", "k_vhgw_stream",
                                                "k_cube2"};
  double ms = 0.0, by = 0.0;
  int64_t n = 0;
  timer_read(&ms, &n, &by);
  TimerState& t = timer_state();
  std::lock_guard<std::mutex> lock(t.mu);
  std::string js = "{";
  for (int k = 0; k < kTagCount; ++k) {
    const TimerState::Total& tt = t.per_tag[k];
    char buf[256];
    std::snprintf(buf, sizeof buf, "%s\"%s\": {\"ms\": %.6f, \"launches\": %lld, \"bytes\": %.1f}", k ? ", " : "",
                  names[k], tt.ms, static_cast<long long>(tt.launches), tt.bytes);
    js += buf;
    t.per_tag[k] = TimerState::Total{};
  }
  return js + "}";
}
void mirror_note(const void* dev, const void* host, size_t bytes) {
  g_mirrors.push_back({static_cast<const char*>(dev), static_cast<const char*>(host), bytes});
}
void mirror_clear() { g_mirrors.clear(); }
const void* mirror_host(const void* dev, size_t bytes) {
  const char* p = static_cast<const char*>(dev);
  for (auto it = g_mirrors.rbegin(); it != g_mirrors.rend(); ++it)
    if (p >= it->dev && p + bytes <= it->dev + it->bytes) return it->host + (p - it->dev);
  return nullptr;
}
int64_t mirror_active(const int32_t* win, int64_t n) {
  const int32_t* h = static_cast<const int32_t*>(mirror_host(win, static_cast<size_t>(n) * sizeof(int32_t)));
  if (h == nullptr) return -1;
  int64_t c = 0;
  for (int64_t b = 0; b < n; ++b) c += h[b] != 0 ? 1 : 0;
  return c;
}

MethodCtl method_ctl(const Geom& g, bool f32) {
  MethodCtl c;
  std::memset(&c, 0, sizeof c);
  c.n_max = 1;
  for (int a = 0; a < g.dim; ++a) c.n_max = std::max(c.n_max, g.n[a]);
  c.count = g.count;
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

namespace {

constexpr double kRungRatio = 1.4142135623730951;
constexpr int kMaxG2Rungs = 32;

// ------------------------------------------------------------------------------------------------
// Device memory. Blocks come from a per-device cache: a released block first waits for the stream
// it was used on, then goes back to the cache for the next call of any thread. (hipFree waits for
// the whole device, which would serialize the calls that other threads run on their own streams.)
// Free blocks are kept up to a quarter of the device's memory; an allocation that fails empties the
// cache and tries once more.
class BlockCache {
 public:
  static BlockCache& get() {
    static BlockCache* c = new BlockCache();  // never destroyed: blocks may be released at exit
    return *c;
  }
  // a block of at least `bytes` (its size in *size)
  void* take(int dev, size_t bytes, size_t* size) {
    const size_t want = round(bytes);
    {
      std::lock_guard<std::mutex> lock(mu_);
      auto& fl = free_[dev];
      auto it = fl.lower_bound(want);
      if (it != fl.end() && it->first <= want + want / 4) {
        void* p = it->second;
        *size = it->first;
        held_[dev] -= it->first;
        fl.erase(it);
        return p;
      }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e == hipErrorOutOfMemory) {
      (void)hipGetLastError();
      trim(dev);
      e = hipMalloc(&p, want);
    }
    if (e == hipErrorOutOfMemory) {
      (void)hipGetLastError();
      throw OutOfDeviceMemory("eda: the device memory does not hold the call's workspace (" +
                              std::to_string(want >> 20) + " MiB more)");
    }
    EDA_HIP(e);
    *size = want;
    return p;
  }
  void put(int dev, void* p, size_t size, hipStream_t s) {
    if (s == nullptr || hipStreamSynchronize(s) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipFree(p);
      return;
    }
    {
      std::lock_guard<std::mutex> lock(mu_);
      if (limit_.count(dev) == 0) {
        size_t fr = 0, total = 0;
        limit_[dev] = hipMemGetInfo(&fr, &total) == hipSuccess ? total / 4 : 0;
      }
      if (held_[dev] + size <= limit_[dev]) {
        free_[dev].emplace(size, p);
        held_[dev] += size;
        return;
      }
    }
    (void)hipFree(p);
  }
  // free bytes the cache holds on a device (they count as available to a call)
  size_t held(int dev) {
    std::lock_guard<std::mutex> lock(mu_);
    return held_[dev];
  }
  void trim(int dev) {
    std::multimap<size_t, void*> fl;
    {
      std::lock_guard<std::mutex> lock(mu_);
      fl.swap(free_[dev]);
      held_[dev] = 0;
    }
    for (auto& kv : fl) (void)hipFree(kv.second);
  }

 private:
  // 64 KiB steps below 16 MiB, 2 MiB steps above
  static size_t round(size_t b) {
    const size_t step = b < (size_t{16} << 20) ? (size_t{64} << 10) : (size_t{2} << 20);
    return (b + step - 1) / step * step;
  }
  std::mutex mu_;
  std::map<int, std::multimap<size_t, void*>> free_;
  std::map<int, size_t> held_, limit_;
};

// A block of device memory used on one stream (the stream is waited for before the block is reused).
struct DevMem {
  void* p = nullptr;
  size_t size = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  DevMem() = default;
  DevMem(size_t bytes, hipStream_t s) : stream(s) {
    if (bytes) {
      EDA_HIP(hipGetDevice(&device));
      p = BlockCache::get().take(device, bytes, &size);
    }
  }
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  DevMem(DevMem&& o) noexcept { swap(o); }
  DevMem& operator=(DevMem&& o) noexcept {
    swap(o);
    return *this;
  }
  void swap(DevMem& o) noexcept {
    std::swap(p, o.p);
    std::swap(size, o.size);
    std::swap(device, o.device);
    std::swap(stream, o.stream);
  }
  ~DevMem() {
    if (p) BlockCache::get().put(device, p, size, stream);
  }
};

}  // namespace

TapeBuf::TapeBuf(size_t bytes) {
  if (bytes == 0) return;
  const hipError_t e = hipMalloc(&p, bytes);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    throw OutOfDeviceMemory("eda: the device memory does not hold the tape (" + std::to_string(bytes >> 20) +
                            " MiB more)");
  }
  EDA_HIP(e);
}
TapeBuf::~TapeBuf() {
  if (p) (void)hipFree(p);
}

namespace {

// A bump allocator over one device allocation.
struct Carve {
  char* base;
  size_t off = 0;
  template <class V>
  V* take(size_t n) {
    off = (off + 255) / 256 * 256;
    V* p = reinterpret_cast<V*>(base + off);
    off += n * sizeof(V);
    return p;
  }
};

// The parameter arena: host values copied through pinned memory into device memory.
struct Uploader;
thread_local Uploader* t_staged = nullptr;  // the uploader with staged, not yet copied, bytes

// Per-step parameter arrays: put() stages them in pinned host memory and returns their device
// address; the staged bytes go to the device as one copy when the next kernel or device copy is
// issued (flush_uploads), so that a burst of uploads costs one copy, not one each.
struct Uploader {
  char* host = nullptr;
  DevMem dev;
  size_t cap = 0, off = 0;
  size_t pbeg = 0, pend = 0;  // the staged range [pbeg, pend)
  hipStream_t s = nullptr;
  void flush() {
    if (pend > pbeg) EDA_HIP(hipMemcpyAsync(static_cast<char*>(dev.p) + pbeg, host + pbeg, pend - pbeg,
                                           hipMemcpyHostToDevice, s));
    pbeg = pend = 0;
    if (t_staged == this) t_staged = nullptr;
  }
  void init(size_t bytes, hipStream_t stream) {
    cap = bytes;
    s = stream;
    EDA_HIP(hipHostMalloc(reinterpret_cast<void**>(&host), cap, hipHostMallocDefault));
    dev = DevMem(cap, stream);
  }
  ~Uploader() {
    if (t_staged == this) t_staged = nullptr;
    if (host) (void)hipHostFree(host);
  }
  template <class V>
  V* put(const V* data, size_t n) {
    const size_t bytes = n * sizeof(V);
    const size_t a = (off + 255) / 256 * 256;
    if (a + bytes > cap) throw std::logic_error("eda: the parameter arena is full");
    if (t_staged != nullptr && t_staged != this) t_staged->flush();
    std::memcpy(host + a, data, bytes);
    char* d = static_cast<char*>(dev.p) + a;
    if (pend == pbeg) pbeg = a;
    pend = a + bytes;
    t_staged = this;
    off = a + bytes;
    mirror_note(d, host + a, bytes);  // the launches read the host copy (timer bytes, window classes)
    return reinterpret_cast<V*>(d);
  }
  template <class V>
  V* put(const std::vector<V>& v) {
    return put(v.data(), v.size());
  }
  void reset() {  // only after the stream is synchronized and no upload is read again
    flush();
    off = 0;
    mirror_clear();
  }
  void reserve(size_t bytes) {
    if (bytes <= cap) return;
    flush();
    EDA_HIP(hipStreamSynchronize(s));
    if (host) EDA_HIP(hipHostFree(host));
    host = nullptr;
    init(bytes, s);
  }
};

}  // namespace

void flush_uploads() {
  if (t_staged != nullptr) t_staged->flush();
}

namespace {

// The per-(thread, device) context: the stream and the parameter arena.
struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  Uploader up;
  explicit Context(int dev) : device(dev) {
    EDA_HIP(hipSetDevice(dev));
    EDA_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    up.init(size_t{4} << 20, stream);
  }
};
// Contexts are pooled per device: a thread takes one at its first call on a device and gives it
// back when it ends, so that a pool of worker threads reuses the streams and every stream that a
// released block names stays valid for the life of the process.
struct ContextPool {
  std::mutex mu;
  std::map<int, std::vector<Context*>> idle;
  static ContextPool& get() {
    static ContextPool* p = new ContextPool();  // never destroyed: threads give contexts back at exit
    return *p;
  }
  Context* take(int device) {
    {
      std::lock_guard<std::mutex> lock(mu);
      auto& v = idle[device];
      if (!v.empty()) {
        Context* c = v.back();
        v.pop_back();
        return c;
      }
    }
    return new Context(device);
  }
  void give(Context* c) {
    std::lock_guard<std::mutex> lock(mu);
    idle[c->device].push_back(c);
  }
};
Context& context(int device) {
  struct Held {
    std::map<int, Context*> m;
    ~Held() {
      for (auto& kv : m) ContextPool::get().give(kv.second);
    }
  };
  thread_local Held held;
  auto it = held.m.find(device);
  if (it == held.m.end()) it = held.m.emplace(device, ContextPool::get().take(device)).first;
  EDA_HIP(hipSetDevice(device));
  return *it->second;
}

// ------------------------------------------------------------------------------------------------
// G2's candidate rungs and confirmation machine (reference driver.cpp:59-116).
struct G2Res {
  int outcome = 0;  // 0 none, 1 accepted, 2 stop
  int32_t s = 0;
  int64_t n_p = 0;
  double p = 0.0;
};
struct RungMachine {
  std::vector<int32_t> rungs;
  int32_t i = 0;
  bool pending = false, done = true, accepted = false;
  int32_t acc_index = -1;
  G2Res cand{};
  int32_t next_rung() const {
    if (done) return 0;
    if (i < static_cast<int32_t>(rungs.size())) return rungs[static_cast<size_t>(i)];
    return odd_window(static_cast<double>(rungs.back()) * kRungRatio + 1.0);
  }
  void feed(const G2Res& o) {
    if (pending) {
      if (o.outcome == 1) {
        done = accepted = true;
        acc_index = i - 1;
        return;
      }
      if (o.outcome == 2) {
        done = true;
        return;
      }
      pending = false;
      ++i;
    } else {
      if (o.outcome == 2) {
        done = true;
        return;
      }
      if (o.outcome == 1) {
        pending = true;
        cand = o;
        cand.s = rungs[static_cast<size_t>(i)];
      }
      ++i;
    }
    if (i >= static_cast<int32_t>(rungs.size()) && !pending) done = true;
  }
};
RungMachine make_rung_machine(int32_t w, int32_t w_prev, double jump) {
  RungMachine m;
  const int32_t s0 = std::max(3, w_prev + 2);
  for (int32_t j = 0;; ++j) {
    const int32_t sv = odd_window(static_cast<double>(s0) * std::pow(kRungRatio, static_cast<double>(j)));
    if (static_cast<double>(sv) * jump > static_cast<double>(w)) break;
    if (m.rungs.empty() || sv != m.rungs.back()) m.rungs.push_back(sv);
    if (static_cast<int32_t>(m.rungs.size()) > kMaxG2Rungs) throw std::logic_error("eda: too many G2 rungs");
  }
  m.done = m.rungs.empty();
  return m;
}

// The rung ladder (reference element.cu:343-370): in double with the C library's log and pow.
std::vector<int32_t> rung_ladder(double d_min, double d_max, int32_t w, int32_t n_max) {
  if (!(std::isfinite(d_min) && std::isfinite(d_max))) throw std::logic_error("eda: rung_ladder: non-finite d");
  volatile double qv = kRungRatio;
  const double q = qv;
  const double lq = std::log(q);
  const double wd = static_cast<double>(w);
  const double jl = std::floor(std::log((d_min > 1.0 ? d_min : 1.0) / wd) / lq + 1e-9);
  const double jh = std::ceil(std::log((d_max > 1.0 ? d_max : 1.0) / wd) / lq - 1e-9);
  if (!(jh - jl <= 4096.0)) throw std::logic_error("eda: rung_ladder: the ladder spans too many rungs");
  const int64_t cap = 2 * static_cast<int64_t>(n_max) + 1;
  std::set<int64_t> sides;
  for (int64_t j = static_cast<int64_t>(jl); j <= static_cast<int64_t>(jh); ++j) {
    const double v = wd * std::pow(q, static_cast<double>(j));
    const int64_t side = v >= static_cast<double>(cap) + 1.0 ? cap : std::min<int64_t>(odd_window(v), cap);
    sides.insert(side);
  }
  if (sides.size() > static_cast<size_t>(kMaxRungs)) throw std::logic_error("eda: the ladder exceeds 32 rungs");
  return std::vector<int32_t>(sides.begin(), sides.end());
}

int32_t mode_bound(const MethodCtl& c) {  // reference driver.cpp:128-140
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

template <class T>
double tie_tol(double tau, double amax) {
  return static_cast<double>(static_cast<T>(tau) * static_cast<T>(amax));
}

// ------------------------------------------------------------------------------------------------
// The rows of a call.
struct Run {
  int device = 0;
  int dtype = 0;
  int dim = 1;
  int64_t n[3] = {1, 1, 1};
  int64_t batch = 0, count = 0;
  int32_t cap = 0;
  int64_t chunk = 0;
  DevMem rows;    // (batch, cap, count) in the compute dtype
  DevMem orient;  // with the orientation kept: (batch, cap - 1, dim + 1, count), n then c of each mode
  bool kept_orient = false;
  std::vector<int32_t> n_imfs, error, windows, sifts;
  std::vector<double> spacing, energy;
  int32_t P = 1, K = 0;
  std::unique_ptr<Tape> tape;  // with eda_params.keep_tape
};

// Per-element host state of a chunk.
struct Rec {
  int status = 0;  // 0 active, 1 done, 2 error
  int32_t err = 0;
  int32_t k = 0, w_prev = 1, w = 0, W = 0, p_odd = 0, kexp = 0, n_imfs = 0, n_sifts = 0;
  bool fine = false, sifting = false;
  double amax = 0.0, tol = 0.0, r_inf = 0.0, sigma_x = 0.0, spacing = 0.0, jump = 0.0, g3 = 1.0, wc = 0.0;
  int32_t Wd = 0, Wl = 0;
};

// One sift's update with the guard maxima (reference sift.cu:160-199): hn = h - a mh, and per
// element the integer maxima of the bits of |r - hn| and |hn| (non-negative doubles order as
// integers), one atomic per wave after a wave64 reduction over all lanes.
template <class T>
__global__ void __launch_bounds__(kThreads) k_sift_step(int64_t count, const int32_t* __restrict__ on,
                                                        const T* __restrict__ a, const T* __restrict__ Mn,
                                                        const T* __restrict__ h, const T* __restrict__ r,
                                                        T* __restrict__ hn, uint64_t* __restrict__ gmax) {
  const int64_t b = blockIdx.y;
  if (on[b] == 0) return;
  uint64_t mr = 0, mh = 0;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kThreads) + threadIdx.x; i < count;
       i += static_cast<int64_t>(gridDim.x) * kThreads) {
    const int64_t at = b * count + i;
    const T m = a[at] * Mn[at];
    const T xv = h[at] - m;
    hn[at] = xv;
    const uint64_t br = abs_bits(r[at] - xv);
    const uint64_t bh = abs_bits(xv);
    mr = br > mr ? br : mr;
    mh = bh > mh ? bh : mh;
  }
  __shared__ uint64_t red[2][kThreads / 64];
  mr = wave_max_u64(mr);
  mh = wave_max_u64(mh);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = mr;
    red[1][threadIdx.x >> 6] = mh;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kThreads / 64; ++w) {
      mr = red[0][w] > mr ? red[0][w] : mr;
      mh = red[1][w] > mh ? red[1][w] : mh;
    }
    if (mr) atomicMax(reinterpret_cast<unsigned long long*>(&gmax[2 * b]), static_cast<unsigned long long>(mr));
    if (mh) atomicMax(reinterpret_cast<unsigned long long*>(&gmax[2 * b + 1]), static_cast<unsigned long long>(mh));
  }
}

// ---- device input: the strided gather and the cast to the compute dtype (reference
// cpp/src/field/convert.hpp:33-49, convert.cu; describe_device_input decompose.cpp:25-82) ----------
size_t source_size(int32_t source, int32_t dtype) {
  switch (source) {
    case EDA_SRC_SAME: return dtype == EDA_F32 ? 4 : 8;
    case EDA_SRC_BOOL: case EDA_SRC_I8: case EDA_SRC_U8: return 1;
    case EDA_SRC_I16: case EDA_SRC_U16: case EDA_SRC_F16: case EDA_SRC_BF16: return 2;
    case EDA_SRC_I32: case EDA_SRC_U32: case EDA_SRC_F32: return 4;
    case EDA_SRC_I64: case EDA_SRC_U64: case EDA_SRC_F64: return 8;
    default: throw std::invalid_argument("eda: unknown source element type " + std::to_string(source));
  }
}

// The batch offset of every element b0 .. b0 + nb - 1 of device input, in source elements.
std::vector<int64_t> batch_offsets(const eda_array_desc& x, int64_t b0, int64_t nb) {
  std::vector<int64_t> off(static_cast<size_t>(nb));
  if (x.batch_axes <= 1) {
    for (int64_t b = 0; b < nb; ++b) off[b] = (b0 + b) * x.batch_stride;
    return off;
  }
  for (int64_t b = 0; b < nb; ++b) {
    int64_t rest = b0 + b, o = 0;
    for (int k = x.batch_axes - 1; k >= 0; --k) {
      const int64_t e = x.batch_shape[k];
      o += (rest % e) * x.batch_strides[k];
      rest /= e;
    }
    off[b] = o;
  }
  return off;
}

struct Bf16 {
  uint16_t bits;
};
struct Bool8 {
  uint8_t v;
};
template <class T, class S>
__device__ __forceinline__ T cast_in(S v) {
  if constexpr (std::is_same_v<S, Bool8>) return v.v != 0 ? T(1) : T(0);
  else if constexpr (std::is_same_v<S, Bf16>) return static_cast<T>(__uint_as_float(static_cast<uint32_t>(v.bits) << 16));
  else return static_cast<T>(v);
}

template <class T, class S>
void gather_cast(hipStream_t s, const Geom& g, const S* xp, const int64_t* boff, const int64_t* st, T* dst) {
  const int64_t s1 = st[0], s2 = st[1], s3 = st[2];
  each(s, g, nullptr, [=] __device__(int64_t b, int64_t i) {
    int64_t off = boff[b] + ((i / g.st[0]) % g.n[0]) * s1;
    if (g.dim >= 2) off += ((i / g.st[1]) % g.n[1]) * s2;
    if (g.dim >= 3) off += ((i / g.st[2]) % g.n[2]) * s3;
    dst[b * g.count + i] = cast_in<T, S>(xp[off]);
  });
}

// dst (C-contiguous (g.batch, count)) <- elements b0 .. b0 + g.batch - 1 of x: host input copied,
// device input gathered through its strides and cast.
template <class T>
void fill_input(Context& ctx, const Geom& g, const eda_array_desc& x, int64_t b0, T* dst) {
  hipStream_t s = ctx.stream;
  if (x.device < 0) {
    EDA_HIP(hipMemcpyAsync(dst, static_cast<const T*>(x.data) + b0 * g.count,
                           static_cast<size_t>(g.batch * g.count) * sizeof(T), hipMemcpyHostToDevice, s));
    return;
  }
  const int64_t* boff = ctx.up.put(batch_offsets(x, b0, g.batch));
  const int64_t* st = x.strides;
  const void* p = x.data;
  const int32_t src = x.source == EDA_SRC_SAME ? (x.dtype == EDA_F32 ? EDA_SRC_F32 : EDA_SRC_F64) : x.source;
  switch (src) {
    case EDA_SRC_BOOL: gather_cast<T>(s, g, static_cast<const Bool8*>(p), boff, st, dst); break;
    case EDA_SRC_I8: gather_cast<T>(s, g, static_cast<const int8_t*>(p), boff, st, dst); break;
    case EDA_SRC_I16: gather_cast<T>(s, g, static_cast<const int16_t*>(p), boff, st, dst); break;
    case EDA_SRC_I32: gather_cast<T>(s, g, static_cast<const int32_t*>(p), boff, st, dst); break;
    case EDA_SRC_I64: gather_cast<T>(s, g, static_cast<const int64_t*>(p), boff, st, dst); break;
    case EDA_SRC_U8: gather_cast<T>(s, g, static_cast<const uint8_t*>(p), boff, st, dst); break;
    case EDA_SRC_U16: gather_cast<T>(s, g, static_cast<const uint16_t*>(p), boff, st, dst); break;
    case EDA_SRC_U32: gather_cast<T>(s, g, static_cast<const uint32_t*>(p), boff, st, dst); break;
    case EDA_SRC_U64: gather_cast<T>(s, g, static_cast<const uint64_t*>(p), boff, st, dst); break;
    case EDA_SRC_F16: gather_cast<T>(s, g, static_cast<const _Float16*>(p), boff, st, dst); break;
    case EDA_SRC_BF16: gather_cast<T>(s, g, static_cast<const Bf16*>(p), boff, st, dst); break;
    case EDA_SRC_F32: gather_cast<T>(s, g, static_cast<const float*>(p), boff, st, dst); break;
    case EDA_SRC_F64: gather_cast<T>(s, g, static_cast<const double*>(p), boff, st, dst); break;
    default: throw std::invalid_argument("eda: unknown source element type");
  }
}

template <class T>
class Chunk {
 public:
  Chunk(Context& ctx, const Geom& g, const MethodCtl& c, double tau, bool check_finite, int32_t max_imf, T* rows,
        int32_t cap, T* orient)
      : ctx_(ctx), g_(g), c_(c), s_(ctx.stream), tau_(tau), check_finite_(check_finite), max_imf_(max_imf),
        rows_(rows), cap_(cap), orient_(orient), B_(g.batch) {
    const int64_t F = g.batch * g.count;
    const int D = g.dim;
    const int K = D * (D + 1) / 2;
    nfields_ = 6 + D + 12 + D + 1 + (D >= 2 ? 3 + K : 0);
    const size_t fbytes = static_cast<size_t>(F) * sizeof(T);
    const size_t small = 64 * 1024 + static_cast<size_t>(B_) * (1024 + sizeof(ElemDev) + kMaxRungs * 4) +
                         static_cast<size_t>(orient_scratch_doubles(B_)) * sizeof(double);
    const size_t sel = std::max(noise_ws_bytes(g), select_ws_bytes(g));
    const size_t bytes = (nfields_ + 1) * (fbytes + 256) + 8 * (static_cast<size_t>(F) + 256) +
                         static_cast<size_t>(F) * 4 + 256 + small + sel + elem_stats_ws_bytes(g) + 4096;
    mem_ = DevMem(bytes, s_);
    Carve cv{static_cast<char*>(mem_.p)};
    r_ = cv.take<T>(F);
    L_ = cv.take<T>(F);
    Lp_ = cv.take<T>(F);
    h_ = cv.take<T>(F);
    t_ = cv.take<T>(F);
    o_ = cv.take<T>(static_cast<size_t>(F) * D);
    for (int i = 0; i < 12; ++i) X_[i] = cv.take<T>(F);
    f1_ = cv.take<T>(F);
    nvec_ = cv.take<T>(static_cast<size_t>(F) * D);
    coh_ = cv.take<T>(F);
    if (D >= 2) owork_ = cv.take<T>(static_cast<size_t>(F) * (3 + K));
    flags_ = cv.take<uint8_t>(F);
    ff_ = cv.take<uint8_t>(F);
    P_[0] = cv.take<uint8_t>(F);
    P_[1] = cv.take<uint8_t>(F);
    Pacc_ = cv.take<uint8_t>(F);
    A_ = cv.take<uint8_t>(F);
    lo_ = cv.take<uint8_t>(F);
    aw_ = cv.take<uint8_t>(F);
    labels_ = cv.take<int32_t>(F);
    sel_ws_ = cv.take<char>(sel);
    stats_ws_ = cv.take<char>(elem_stats_ws_bytes(g));
    d_sigma_n_ = cv.take<double>(B_);
    d_e_med_ = cv.take<double>(B_);
    d_wc_ = cv.take<double>(B_);
    d_p_ = cv.take<double>(B_);
    d_energy_ = cv.take<double>(B_);
    d_amax_ = cv.take<uint64_t>(B_);
    d_nonfin_cnt_ = cv.take<uint64_t>(B_);
    d_ext_ = cv.take<uint64_t>(3 * B_);
    d_np_ = cv.take<int64_t>(B_);
    d_nonfinite_ = cv.take<int32_t>(B_);
    d_any_ = cv.take<int32_t>(2 * B_);
    d_dminmax_ = cv.take<uint64_t>(2 * B_);
    d_bits_ = cv.take<uint32_t>(2 * B_);
    d_gmax_ = cv.take<uint64_t>(2 * B_);
    d_el_ = cv.take<ElemDev>(B_);
    d_rung_win_ = cv.take<int32_t>(static_cast<size_t>(kMaxRungs) * B_);
    d_oscr_ = cv.take<double>(static_cast<size_t>(orient_scratch_doubles(B_)));
    if (cv.off > bytes) throw std::logic_error("eda: the workspace layout overflows");
    rec_.assign(static_cast<size_t>(B_), Rec{});
    ctx_.up.reserve((size_t{16} << 20) + static_cast<size_t>(B_) * 16384);
  }

  // Copies the input (device pointer with element strides, or contiguous host memory) into r.
  // Copies elements b0 .. b0 + B - 1 of the call's input into r (host input copied, device input
  // gathered through its strides and cast).
  void start(const eda_array_desc& x, int64_t b0) {
    ctx_.up.reset();
    const Geom g = g_;
    T* r = r_;
    fill_input<T>(ctx_, g, x, b0, r);
    // non-finite samples and max |x|; the exact power-of-two prescale (reference mode.cu:43-58,271-282)
    enqueue_elem_stats<T>(s_, g_, r_, nullptr, nullptr, d_amax_, d_nonfin_cnt_, stats_ws_);
    std::vector<uint64_t> amax(B_), nf(B_);
    d2h(amax.data(), d_amax_, B_);
    d2h(nf.data(), d_nonfin_cnt_, B_);
    sync();
    std::vector<int32_t> kexp(B_, 0);
    for (int64_t b = 0; b < B_; ++b) {
      const double a = bits_d(amax[b]);
      kexp[b] = (a > 0.0 && std::isfinite(a)) ? std::ilogb(a) : 0;
      rec_[b].kexp = kexp[b];
      if (nf[b] != 0 && check_finite_) {
        rec_[b].status = 2;
        rec_[b].err = 1;
      }
    }
    const int32_t* dk = ctx_.up.put(kexp);
    each(s_, g, nullptr, [=] __device__(int64_t b, int64_t i) {
      const int32_t k = dk[b];
      if (k == 0) return;
      const int64_t at = b * g.count + i;
      if constexpr (std::is_same_v<T, float>) r[at] = static_cast<float>(ldexp(static_cast<double>(r[at]), -k));
      else r[at] = ldexp(r[at], -k);
    });
    enqueue_elem_stats<T>(s_, g_, r_, nullptr, nullptr, d_amax_, nullptr, stats_ws_);
    d2h(amax.data(), d_amax_, B_);
    sync();
    for (int64_t b = 0; b < B_; ++b) {
      rec_[b].amax = bits_d(amax[b]);
      rec_[b].tol = tie_tol<T>(tau_, rec_[b].amax);
    }
    // sigma_x: the largest per-axis noise level of the prescaled input
    const int32_t* on = upload_on([](const Rec& e) { return e.status == 0; });
    enqueue_noise_sigma<T>(s_, g_, r_, true, d_sigma_n_, on, sel_ws_);
    std::vector<double> sx(B_);
    d2h(sx.data(), d_sigma_n_, B_);
    sync();
    for (int64_t b = 0; b < B_; ++b) rec_[b].sigma_x = rec_[b].status == 0 ? sx[b] : 0.0;
  }

  // One mode of every active element (reference driver.cpp:423-689); false when none is active.
  bool mode();

  // The residue rows and the unscale (reference field_ops.cu:157-174, mode.cu:284-302).
  void finish(Run& run, int64_t b0) {
    ctx_.up.reset();
    const Geom g = g_;
    std::vector<int32_t> row_of(B_), kx(B_);
    for (int64_t b = 0; b < B_; ++b) {
      row_of[b] = rec_[b].status == 2 ? -1 : rec_[b].n_imfs;
      kx[b] = rec_[b].kexp;
    }
    const int32_t* drow = ctx_.up.put(row_of);
    const int32_t* dk = ctx_.up.put(kx);
    T* rows = rows_;
    T* r = r_;
    const int64_t cap = cap_;
    int32_t* bad = d_nonfinite_;
    EDA_HIP(hipMemsetAsync(bad, 0, sizeof(int32_t) * B_, s_));
    each(s_, g, nullptr, [=] __device__(int64_t b, int64_t i) {
      const int32_t nr = drow[b];
      if (nr < 0) return;
      rows[(b * cap + nr) * g.count + i] = r[b * g.count + i];
    });
    // rows 0 .. nr of element b: samples i < (nr + 1) count of its rows (no index division), over
    // the rows up to the largest nr
    int32_t top = 0;
    for (int64_t b = 0; b < B_; ++b) top = std::max(top, row_of[b]);
    Geom gr = g;
    gr.count = (static_cast<int64_t>(top) + 1) * g.count;
    each(s_, gr, nullptr, [=] __device__(int64_t b, int64_t i) {
      const int32_t k = dk[b];
      const int32_t nr = drow[b];
      if (k == 0 || nr < 0) return;
      if (i >= (static_cast<int64_t>(nr) + 1) * g.count) return;
      const int64_t idx = b * cap * g.count + i;
      T v;
      if constexpr (std::is_same_v<T, float>) v = static_cast<float>(ldexp(static_cast<double>(rows[idx]), k));
      else v = ldexp(rows[idx], k);
      rows[idx] = v;
      if (!isfinite(v)) bad[b] = 1;
    });
    std::vector<int32_t> hb(B_);
    d2h(hb.data(), bad, B_);
    sync();
    for (int64_t b = 0; b < B_; ++b) {
      Rec& e = rec_[b];
      const int64_t ib = b0 + b;
      int32_t err = 0;
      if (e.status == 2) err = e.err ? e.err : 255;
      else if (hb[b]) err = 2;
      else if (e.status != 1) err = 255;
      run.error[ib] = err;
      run.n_imfs[ib] = err ? 0 : e.n_imfs;
      const int32_t hcap = cap_ - 1;
      for (int32_t j = 0; j < hcap; ++j) {
        const size_t h = static_cast<size_t>(ib) * hcap + j;
        const bool keep = err == 0 && j < e.n_imfs;
        run.windows[h] = keep ? hist_w_[b][j] : 0;
        run.sifts[h] = keep ? hist_s_[b][j] : 0;
        run.spacing[h] = keep ? hist_sp_[b][j] : 0.0;
        run.energy[h] = keep ? std::ldexp(hist_e_[b][j], 2 * e.kexp) : 0.0;
      }
    }
  }

  // return_orientation: row n_imfs (the row of the mode this element commits) of orient_ (B, cap - 1,
  // D + 1, count) <- n with the canonical sign (components 0 .. D - 1) and c (component D) of the
  // active elements (reference driver.cpp:553-559, orientation.cu:495-522). The canonical sign
  // negates n when its first component, in axis order, of magnitude above the dtype's default tie
  // tolerance (2^-14 float32, 2^-30 float64) is negative (orientation.cu:405-417).
  void store_orientation(const int32_t* on) {
    std::vector<int32_t> kv(B_, 0);
    for (int64_t b = 0; b < B_; ++b)
      if (rec_[b].status == 0) kv[b] = rec_[b].n_imfs;
    const int32_t* drow = ctx_.up.put(kv);
    const Geom g = g_;
    const int D = g.dim;
    const int64_t B = B_, rows = cap_ - 1;
    const T zero = std::is_same_v<T, float> ? T(0x1p-14) : T(0x1p-30);
    const T* nv = nvec_;
    const T* cv = coh_;
    T* dst = orient_;
    each(s_, g, on, [=] __device__(int64_t b, int64_t i) {
      T v[3];
      int sgn = 0;
      for (int a = 0; a < D; ++a) {
        v[a] = nv[(a * B + b) * g.count + i];
        if (sgn == 0 && fabs(v[a]) > zero) sgn = v[a] > T(0) ? 1 : -1;
      }
      T* o = dst + (b * rows + drow[b]) * (D + 1) * g.count + i;
      for (int a = 0; a < D; ++a) o[a * g.count] = sgn < 0 ? -v[a] : v[a];
      o[D * g.count] = cv[b * g.count + i];
    });
  }

  // Keep the decisions of every mode in t (the adjoint's tape).
  void keep_tape(ChunkTape* t) { tape_ = t; }
  // The per-element results the backward reads: the prescale exponents, the IMF counts, the errors.
  void tape_results(ChunkTape& t, const Run& run, int64_t b0) const {
    t.kexp.assign(B_, 0);
    t.n_imfs.assign(B_, 0);
    t.error.assign(B_, 0);
    for (int64_t b = 0; b < B_; ++b) {
      t.kexp[b] = rec_[b].kexp;
      t.n_imfs[b] = run.n_imfs[b0 + b];
      t.error[b] = run.error[b0 + b];
    }
  }

  void alloc_history() {
    hist_w_.assign(B_, std::vector<int32_t>(cap_ - 1, 0));
    hist_s_.assign(B_, std::vector<int32_t>(cap_ - 1, 0));
    hist_sp_.assign(B_, std::vector<double>(cap_ - 1, 0.0));
    hist_e_.assign(B_, std::vector<double>(cap_ - 1, 0.0));
  }

 private:
  template <class V>
  void d2h(V* dst, const V* src, int64_t n) {
    EDA_HIP(hipMemcpyAsync(dst, src, sizeof(V) * n, hipMemcpyDeviceToHost, s_));
  }
  void sync() { EDA_HIP(hipStreamSynchronize(s_)); }
  template <class P>
  const int32_t* upload_on(P pred) {
    std::vector<int32_t> on(B_);
    for (int64_t b = 0; b < B_; ++b) on[b] = pred(rec_[b]) ? 1 : 0;
    return ctx_.up.put(on);
  }
  // A window array of the elements that satisfy pred (0 elsewhere) and its maximum.
  template <class W, class P>
  const int32_t* upload_win(W winf, P pred, int32_t* maxw) {
    std::vector<int32_t> w(B_, 0);
    int32_t m = 0;
    for (int64_t b = 0; b < B_; ++b) {
      if (!pred(rec_[b])) continue;
      w[b] = winf(rec_[b]);
      m = std::max(m, w[b]);
    }
    *maxw = m;
    return ctx_.up.put(w);
  }
  template <class F, class P>
  const double* upload_f64(F f, P pred) {
    std::vector<double> v(B_, 0.0);
    for (int64_t b = 0; b < B_; ++b)
      if (pred(rec_[b])) v[b] = f(rec_[b]);
    return ctx_.up.put(v);
  }

  void tau_field(const int32_t* on, const int32_t* dWd, int32_t mWd, const int32_t* dWl, int32_t mWl,
                 const double* tol, const double* floor2, T* tau, T* tmp);
  void raw_window(const uint8_t* flags, const int32_t* win, int32_t maxw, const int32_t* on, T* om, T* tmp);
  // The envelopes u, l of h and the box of their mean 0.5 (u + l) into m (the mean comes out of
  // the max / min gather); optionally also hd = 0.5 (u - l), R = max(u - l, 0) and d = h - m.
  void envelope_mean(const T* h, T* m, T* u, T* l, T* hd = nullptr, T* R = nullptr, T* d = nullptr) {
    GatherExtra<T> mm;
    mm.mid = X_[3];
    mm.hdiff = hd;
    mm.pdiff = R;
    elem_op(2, h, u, l, mm);
    GatherExtra<T> bx;
    bx.minuend = h;
    bx.diff = d;
    elem_op(3, mm.mid, m, nullptr, bx);
  }
  // An operator of the element over every rung (the rung scratch grows to the mode's ladder), with
  // the gather's extra fields.
  void elem_op(int op, const T* in, T* out0, T* out1, const GatherExtra<T>& extra = {}) {
    // the rung fields: r0, r1 (K_max fields each), then the max and min side by side (2 K_max)
    const size_t need = static_cast<size_t>(K_max_) * static_cast<size_t>(B_ * g_.count);
    if (need > rung_cap_) {
      EDA_HIP(hipStreamSynchronize(s_));
      rung_mem_ = DevMem();
      rung_mem_ = DevMem(4 * need * sizeof(T), s_);
      rung_cap_ = need;
    }
    T* r0 = static_cast<T*>(rung_mem_.p);
    T* r1 = r0 + rung_cap_;
    T* pairs = r1 + rung_cap_;
    enqueue_element_op<T>(s_, g_, op, d_el_, lo_, t_, o_, d_rung_win_, rung_max_.data(), K_max_, in, out0, out1, r0, r1,
                          pairs, on_op_, extra);
  }

  Context& ctx_;
  Geom g_;
  MethodCtl c_;
  hipStream_t s_;
  double tau_;
  bool check_finite_;
  int32_t max_imf_;
  T* rows_;
  int32_t cap_;
  T* orient_;  // nullptr: the orientation is not kept
  int64_t B_;
  int nfields_ = 0;
  DevMem mem_;
  T *r_, *L_, *Lp_, *h_, *t_, *o_, *nvec_, *coh_, *owork_ = nullptr;
  T* X_[12];
  T* f1_ = nullptr;
  DevMem rung_mem_;
  size_t rung_cap_ = 0;
  uint8_t *flags_, *ff_, *P_[2], *Pacc_, *A_, *lo_, *aw_;
  int32_t* labels_;
  char *sel_ws_, *stats_ws_;
  double *d_sigma_n_, *d_e_med_, *d_wc_, *d_p_, *d_energy_;
  uint64_t *d_amax_, *d_nonfin_cnt_, *d_ext_, *d_dminmax_, *d_gmax_;
  int64_t* d_np_;
  int32_t *d_nonfinite_, *d_any_;
  uint32_t* d_bits_;
  ElemDev* d_el_;
  int32_t* d_rung_win_;
  double* d_oscr_;
  const int32_t* on_op_ = nullptr;
  ChunkTape* tape_ = nullptr;
  std::vector<int32_t> rw_, rw_cur_;
  std::vector<int32_t> rung_max_;
  int K_max_ = 0;
  std::vector<Rec> rec_;
  std::vector<std::vector<int32_t>> hist_w_, hist_s_;
  std::vector<std::vector<double>> hist_sp_, hist_e_;
};

// The threshold field (reference stats_fields.cu:26-69,375-389) into tau; nonfinite flags into
// d_nonfinite_ (the caller zeroes them).
template <class T>
void Chunk<T>::tau_field(const int32_t* on, const int32_t* dWd, int32_t mWd, const int32_t* dWl, int32_t mWl,
                         const double* tol, const double* floor2, T* tau, T* tmp) {
  const Geom g = g_;
  const T* r = r_;
  enqueue_cube<T>(s_, g_, CubeOp::kBox, r_, tmp, nullptr, dWd, mWd);
  each(s_, g, on, [=] __device__(int64_t b, int64_t i) {
    const int64_t at = b * g.count + i;
    const T det = r[at] - tmp[at];
    tau[at] = det * det;
  });
  enqueue_cube<T>(s_, g_, CubeOp::kBox, tau, tau, nullptr, dWl, mWl);
  const double theta = c_.sig_theta, k_noise = c_.sig_noise;
  const double* sigma_n = d_sigma_n_;
  int32_t* nonfinite = d_nonfinite_;
  each(s_, g, on, [=] __device__(int64_t b, int64_t i) {
    const int64_t at = b * g.count + i;
    const T be = tau[at];
    if (!(isfinite(be) && isfinite(sigma_n[b]))) nonfinite[b] = 1;
    const T lev = sqrt(be > T(0) ? be : T(0));
    const double f = k_noise * sigma_n[b];
    const double fl = f > 0.0 ? f : 0.0;
    const double ld = static_cast<double>(lev);
    double t = theta * (ld > fl ? ld : fl);
    const double tl = tol[b];
    t = t > tl ? t : tl;
    if (floor2 != nullptr) t = t > floor2[b] ? t : floor2[b];
    tau[at] = static_cast<T>(t);
  });
}

// The raw local window (reference stats_fields.cu:334-369,425-440).
template <class T>
void Chunk<T>::raw_window(const uint8_t* flags, const int32_t* win, int32_t maxw, const int32_t* on, T* om, T* tmp) {
  const Geom g = g_;
  const T two_kappa = static_cast<T>(2.0 * c_.kappa);
  const T tiny = static_cast<T>(c_.tiny);
  // axis a's flags as 0 / 1 into tmp (for a > 0 in the pass that accumulates axis a - 1's square)
  each(s_, g, on, [=] __device__(int64_t b, int64_t i) {
    const int64_t at = b * g.count + i;
    tmp[at] = (flags[at] & 1u) ? T(1) : T(0);
  });
  for (int a = 0; a < g.dim; ++a) {
    enqueue_cube<T>(s_, g_, CubeOp::kBox, tmp, tmp, nullptr, win, maxw);
    const bool last = a == g.dim - 1;
    each(s_, g, on, [=] __device__(int64_t b, int64_t i) {
      const int64_t at = b * g.count + i;
      const T sq = tmp[at] * tmp[at];
      T acc = a == 0 ? sq : om[at] + sq;
      if (last) {
        const T dens = sqrt(acc);
        acc = dens > T(0) ? two_kappa / (dens > tiny ? dens : tiny) : inf_t<T>();
      } else {
        tmp[at] = ((flags[at] >> (a + 1)) & 1u) ? T(1) : T(0);
      }
      om[at] = acc;
    });
  }
}

template <class T>
bool Chunk<T>::mode() {
  const Geom g = g_;
  const MethodCtl c = c_;
  const int D = g.dim;
  const int64_t B = B_;
  auto active = [](const Rec& e) { return e.status == 0; };
  bool any = false;
  for (const Rec& e : rec_) any = any || e.status == 0;
  if (!any) return false;
  ctx_.up.reset();  // the previous mode ended with a synchronization

  // step 0: mode start (reference mode.cu:95-136); sigma_n
  for (Rec& e : rec_) {
    if (e.status != 0) continue;
    e.r_inf = e.amax;
    e.fine = false;
    e.jump = c.absent_jump;
    e.g3 = 1.0;
    e.wc = INFINITY;
    e.p_odd = 0;
    e.n_sifts = 0;
    e.sifting = true;
    const int32_t w0 = std::max(3, e.w_prev + 2);
    e.Wd = odd_window(c.sig_rho * static_cast<double>(w0));
    e.Wl = odd_window(std::max(c.beta, c.sig_rho) * static_cast<double>(w0));
  }
  const int32_t* on = upload_on(active);
  enqueue_noise_sigma<T>(s_, g_, r_, false, d_sigma_n_, on, sel_ws_);
  T* tau = X_[0];
  T* tmp = X_[1];
  // steps 1-2: the window passes (reference mode.cu:138-187)
  for (int32_t pass = 0; pass < c.sig_passes; ++pass) {
    on = upload_on(active);
    int32_t mWd, mWl;
    const int32_t* dWd = upload_win([](const Rec& e) { return e.Wd; }, active, &mWd);
    const int32_t* dWl = upload_win([](const Rec& e) { return e.Wl; }, active, &mWl);
    const double* dtol = upload_f64([](const Rec& e) { return e.tol; }, active);
    EDA_HIP(hipMemsetAsync(d_nonfinite_, 0, sizeof(int32_t) * B, s_));
    EDA_HIP(hipMemsetAsync(d_ext_, 0, sizeof(uint64_t) * 3 * B, s_));
    tau_field(on, dWd, mWd, dWl, mWl, dtol, nullptr, tau, tmp);
    enqueue_sig_flags<T>(s_, g_, r_, tau, flags_, d_ext_, on, tmp);
    std::vector<uint64_t> ext(3 * B);
    std::vector<int32_t> nf(B);
    d2h(ext.data(), d_ext_, 3 * B);
    d2h(nf.data(), d_nonfinite_, B);
    sync();
    for (int64_t b = 0; b < B; ++b) {
      Rec& e = rec_[b];
      if (e.status != 0) continue;
      if (nf[b]) {
        e.status = 2;
        e.err = 2;
        continue;
      }
      const uint64_t N = static_cast<uint64_t>(g.count);
      uint64_t S2 = 0;
      bool enough = false;
      for (int a = 0; a < D; ++a) {
        const uint64_t ea = ext[3 * b + a];
        S2 += ea * ea;
        if (ea * static_cast<uint64_t>(g.n[a]) >= 3 * N) enough = true;
      }
      if (!enough) {
        e.status = 1;
        continue;
      }
      const double sp = static_cast<double>(N) / std::sqrt(static_cast<double>(S2));
      const double v = (2.0 * c.kappa) * sp;
      if (!(v < 1073741824.0)) {
        e.status = 1;
        continue;
      }
      int32_t w = 2 * static_cast<int32_t>(std::floor(std::max(v, 1.0) / 2.0)) + 1;
      w = std::max(w, 3);
      w = std::max(w, e.w_prev + 2);
      if (w > c.n_max) {
        e.status = 1;
        continue;
      }
      e.spacing = sp;
      e.w = w;
      e.Wd = odd_window(c.sig_rho * static_cast<double>(w));
      e.Wl = odd_window(std::max(c.beta, c.sig_rho) * static_cast<double>(w));
    }
  }
  any = false;
  for (const Rec& e : rec_) any = any || e.status == 0;
  if (!any) return false;

  // step 3: G2 (reference driver.cpp:445-509)
  std::vector<RungMachine> mach(static_cast<size_t>(B));
  bool any_g2 = false;
  for (int64_t b = 0; b < B; ++b) {
    const Rec& e = rec_[b];
    if (e.status != 0) continue;
    const int32_t s0 = std::max(3, e.w_prev + 2);
    if (c.fine_jump * static_cast<double>(s0) <= static_cast<double>(e.w)) {
      mach[b] = make_rung_machine(e.w, e.w_prev, c.fine_jump);
      any_g2 = any_g2 || !mach[b].done;
    }
  }
  const uint8_t* ffp = flags_;
  if (any_g2) {
    auto in_g2 = [&](int64_t b) { return rec_[b].status == 0 && !mach[b].done; };
    std::vector<int32_t> ong(B);
    for (int64_t b = 0; b < B; ++b) ong[b] = in_g2(b) ? 1 : 0;
    std::vector<int32_t> g2nf(B, 0);
    if (c.fine_theta > 0) {
      const int32_t* don = ctx_.up.put(ong);
      std::vector<int32_t> wd(B, 0), wl(B, 0);
      std::vector<double> tol(B, 0.0), fl2(B, 0.0);
      int32_t mWd = 0, mWl = 0;
      for (int64_t b = 0; b < B; ++b) {
        if (!ong[b]) continue;
        wd[b] = rec_[b].Wd;
        wl[b] = rec_[b].Wl;
        mWd = std::max(mWd, wd[b]);
        mWl = std::max(mWl, wl[b]);
        tol[b] = rec_[b].tol;
        fl2[b] = c.fine_theta * rec_[b].sigma_x;
      }
      const int32_t* dWd = ctx_.up.put(wd);
      const int32_t* dWl = ctx_.up.put(wl);
      const double* dtol = ctx_.up.put(tol);
      const double* dfl2 = ctx_.up.put(fl2);
      EDA_HIP(hipMemsetAsync(d_nonfinite_, 0, sizeof(int32_t) * B, s_));
      EDA_HIP(hipMemsetAsync(d_ext_, 0, sizeof(uint64_t) * 3 * B, s_));
      tau_field(don, dWd, mWd, dWl, mWl, dtol, dfl2, tau, tmp);
      enqueue_sig_flags<T>(s_, g_, r_, tau, ff_, d_ext_, don, tmp);
      d2h(g2nf.data(), d_nonfinite_, B);
      ffp = ff_;
    }
    T* om = X_[2];
    T* t1 = X_[3];
    for (int32_t step = 0;; ++step) {
      std::vector<int32_t> sr(B, 0);
      bool anyr = false;
      for (int64_t b = 0; b < B; ++b) {
        if (rec_[b].status != 0 || mach[b].done) continue;
        sr[b] = mach[b].next_rung();
        anyr = true;
      }
      if (!anyr) break;
      if (step > kMaxG2Rungs + 1) throw std::logic_error("eda: the G2 rung loop did not end");
      uint8_t* P = P_[step & 1];
      // one rung test per element (reference stats_rungs.cu:16-87,126-159)
      std::vector<int32_t> wb(B, 0), wcore(B, 0), wr(B, 0);
      int32_t mwb = 0, mwc = 0, mwr = 0;
      std::vector<double> lo_s(B, 0.0);
      for (int64_t b = 0; b < B; ++b) {
        if (!sr[b]) continue;
        const double sd = static_cast<double>(sr[b]);
        wb[b] = odd_window(c.fine_beta * sd);
        wcore[b] = odd_window(c.fine_core * sd);
        wr[b] = odd_window((2.0 * c.absent_reach) * static_cast<double>(wb[b]));
        mwb = std::max(mwb, wb[b]);
        mwc = std::max(mwc, wcore[b]);
        mwr = std::max(mwr, wr[b]);
        lo_s[b] = c.fine_lo * sd;
      }
      const int32_t* don = ctx_.up.put(sr);
      const int32_t* dwb = ctx_.up.put(wb);
      const int32_t* dwc = ctx_.up.put(wcore);
      const int32_t* dwr = ctx_.up.put(wr);
      const double* dlo = ctx_.up.put(lo_s);
      raw_window(ffp, dwb, mwb, don, om, tmp);
      each(s_, g, don, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        const bool in = static_cast<double>(om[at]) < dlo[b];
        P[at] = in ? 1 : 0;
        t1[at] = in ? T(1) : T(0);
      });
      enqueue_cube<T>(s_, g_, CubeOp::kMin, t1, t1, nullptr, dwc, mwc);
      int32_t* any_core = d_any_;
      int32_t* any_jump = d_any_ + B;
      EDA_HIP(hipMemsetAsync(d_any_, 0, sizeof(int32_t) * 2 * B, s_));
      each(s_, g, don, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        if (t1[at] > T(0.5)) any_core[b] = 1;
        t1[at] = (P[at] & 1u) ? T(1) : T(0);
      });
      Selection<T> sel{};
      sel.src = SelSrc::kField;
      sel.field = om;
      sel.mask = P;
      sel.bits = 1;
      enqueue_select<T>(s_, g_, sel, 1, c.fine_pop, d_p_, d_np_, don, sel_ws_);
      enqueue_cube<T>(s_, g_, CubeOp::kMax, t1, t1, nullptr, dwr, mwr);
      const double reset = c.reset;
      const double* dp = d_p_;
      each(s_, g, don, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        if (static_cast<double>(om[at]) > reset * dp[b] && t1[at] > T(0.5)) any_jump[b] = 1;
      });
      std::vector<double> hp(B);
      std::vector<int64_t> hnp(B);
      std::vector<int32_t> hany(2 * B);
      d2h(hp.data(), d_p_, B);
      d2h(hnp.data(), d_np_, B);
      d2h(hany.data(), d_any_, 2 * B);
      sync();
      for (int64_t b = 0; b < B; ++b) {
        if (!sr[b]) continue;
        G2Res res;
        res.s = sr[b];
        res.n_p = hnp[b];
        res.p = hp[b];
        if (res.n_p == 0 || !hany[b]) res.outcome = 0;
        else if (c.fine_jump * res.p > static_cast<double>(rec_[b].w)) res.outcome = 2;
        else res.outcome = hany[B + b] ? 1 : 0;
        mach[b].feed(res);
      }
    }
    bool any_acc = false;
    std::vector<int32_t> acc_src(B, 0);  // 1 + the mask buffer of an accepted population, 0: none
    for (int64_t b = 0; b < B; ++b) {
      if (rec_[b].status != 0 || !mach[b].accepted) continue;
      Rec& e = rec_[b];
      e.fine = true;
      e.p_odd = odd_window(mach[b].cand.p);
      acc_src[b] = 1 + (mach[b].acc_index & 1);
      any_acc = true;
    }
    if (any_acc) {
      // the accepted populations into Pacc in one launch
      const int32_t* dsrc = ctx_.up.put(acc_src);
      uint8_t* const pacc = Pacc_;
      const uint8_t* const p0 = P_[0];
      const uint8_t* const p1 = P_[1];
      const int64_t cnt = g.count;
      each(s_, g, dsrc, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * cnt + i;
        pacc[at] = dsrc[b] == 2 ? p1[at] : p0[at];
      });
      // G3's coarse scale (reference stats_rungs.cu:90-121,161-187)
      std::vector<int32_t> wc(B, 0), wrr(B, 0), fon(B, 0);
      int32_t mwc = 0, mwr = 0;
      for (int64_t b = 0; b < B; ++b) {
        const Rec& e = rec_[b];
        if (e.status != 0 || !e.fine) continue;
        fon[b] = 1;
        wc[b] = odd_window(c.beta * static_cast<double>(e.w));
        const int32_t wf = odd_window(c.fine_beta * static_cast<double>(e.p_odd));
        wrr[b] = odd_window((2.0 * c.absent_reach) * static_cast<double>(wf));
        mwc = std::max(mwc, wc[b]);
        mwr = std::max(mwr, wrr[b]);
      }
      const int32_t* don = ctx_.up.put(fon);
      const int32_t* dwc = ctx_.up.put(wc);
      const int32_t* dwr = ctx_.up.put(wrr);
      raw_window(ffp, dwc, mwc, don, om, tmp);
      const uint8_t* Pa = Pacc_;
      each(s_, g, don, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        t1[at] = (Pa[at] & 1u) ? T(1) : T(0);
      });
      enqueue_cube<T>(s_, g_, CubeOp::kMax, t1, t1, nullptr, dwr, mwr);
      uint8_t* selm = aw_;
      each(s_, g, don, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        selm[at] = (t1[at] < T(0.5) && om[at] < inf_t<T>()) ? 1 : 0;
      });
      Selection<T> sel{};
      sel.src = SelSrc::kField;
      sel.field = om;
      sel.mask = selm;
      sel.bits = 1;
      enqueue_select<T>(s_, g_, sel, 0, 0.5, d_wc_, d_np_, don, sel_ws_);
      std::vector<double> hwc(B);
      std::vector<int64_t> hn(B);
      d2h(hwc.data(), d_wc_, B);
      d2h(hn.data(), d_np_, B);
      sync();
      for (int64_t b = 0; b < B; ++b)
        if (fon[b]) rec_[b].wc = hn[b] == 0 ? INFINITY : hwc[b];
    } else {
      sync();
    }
    for (int64_t b = 0; b < B; ++b)
      if (ong[b] && g2nf[b] && rec_[b].status == 0) {
        rec_[b].status = 2;
        rec_[b].err = 2;
      }
  }

  // step 4: the window (reference mode.cu:189-223)
  for (Rec& e : rec_) {
    if (e.status != 0) continue;
    int32_t w = e.w;
    if (e.fine) {
      w = std::max(std::max(e.p_odd, e.w_prev + 2), 3);
      e.jump = 1.0;
    }
    if (c.growth > 1.0 && e.k >= c.growth_from) {
      w = std::max(w, odd_window(c.growth * static_cast<double>(e.w_prev)));
      if (w > c.n_max) {
        e.status = 1;
        continue;
      }
    }
    e.w = w;
    e.W = odd_window(c.beta * static_cast<double>(w));
    if (e.fine && std::isfinite(e.wc)) {
      const double q = std::sqrt(e.wc / static_cast<double>(e.p_odd));
      double f = (c.split * q) / 2.0;
      e.g3 = f < 1.0 ? f : 1.0;
    }
  }
  any = false;
  for (const Rec& e : rec_) any = any || e.status == 0;
  if (!any) return false;

  // step 5: the raw window, the scale field, the floor, G3's scaling (reference stats_scale.cu)
  on = upload_on(active);
  int32_t mW;
  const int32_t* dW = upload_win([](const Rec& e) { return e.W; }, active, &mW);
  T* om = X_[0];
  raw_window(flags_, dW, mW, on, om, tmp);
  {
    std::vector<int32_t> wv(B, 0), wn(B, 0);
    std::vector<double> jump(B, 1.0);
    int32_t mwn = 0;
    for (int64_t b = 0; b < B; ++b) {
      const Rec& e = rec_[b];
      if (e.status != 0) continue;
      wv[b] = e.w;
      jump[b] = e.jump;
      wn[b] = odd_window((2.0 * c.absent_reach) * static_cast<double>(e.W));
      mwn = std::max(mwn, wn[b]);
    }
    const int32_t* dw = ctx_.up.put(wv);
    const int32_t* dwn = ctx_.up.put(wn);
    const double* djump = ctx_.up.put(jump);
    T* fine = X_[2];
    uint8_t* A = A_;
    uint8_t* aw = aw_;
    const double reset = c.reset, lo = c.absent_lo;
    each(s_, g, on, [=] __device__(int64_t b, int64_t i) {
      const int64_t at = b * g.count + i;
      const double w = static_cast<double>(dw[b]);
      const double o = static_cast<double>(om[at]);
      const bool a0 = o > reset * w;
      const bool awv = o > lo * w;
      A[at] = a0 ? 4 : 0;
      aw[at] = awv ? 1 : 0;
      fine[at] = o < (lo * w) / djump[b] ? T(1) : T(0);
    });
    enqueue_cube<T>(s_, g_, CubeOp::kMax, fine, fine, nullptr, dwn, mwn);  // near
    enqueue_ccl(s_, g_, aw_, 1, labels_, on);
    uint8_t* marks = P_[0];  // free after G2 (its accepted population lives in Pacc)
    const int32_t* labels = labels_;
    each(s_, g, on, [=] __device__(int64_t b, int64_t i) { marks[b * g.count + i] = 0; });
    each(s_, g, on, [=] __device__(int64_t b, int64_t i) {
      const int64_t at = b * g.count + i;
      if (aw[at] && fine[at] > T(0.5)) marks[b * g.count + labels[at]] = 1;
    });
    T* L = L_;
    const double absent = c.absent, reset2 = c.reset, c_lo = c.c_lo, c_hi = c.c_hi;
    each(s_, g, on, [=] __device__(int64_t b, int64_t i) {
      const int64_t at = b * g.count + i;
      const int32_t lab = labels[at];
      const bool a = (A[at] & 4u) && lab >= 0 && marks[b * g.count + lab];
      A[at] = a ? 1 : 0;
      const double w = static_cast<double>(dw[b]);
      double v;
      if (a) {
        v = absent;
      } else {
        const double o = static_cast<double>(om[at]);
        const double ow = o > reset2 * w ? w : o;
        const double lo_ = w / c_lo;
        const double hi = c_hi * w;
        v = ow > lo_ ? ow : lo_;
        v = v < hi ? v : hi;
      }
      L[at] = log(static_cast<T>(v));
    });
    enqueue_cube<T>(s_, g_, CubeOp::kBox, L_, L_, nullptr, dW, mW);
    each(s_, g, on, [=] __device__(int64_t b, int64_t i) {
      const int64_t at = b * g.count + i;
      L[at] = exp(L[at]);
    });
    // the per-voxel floor for modes k >= 1 (reference stats_scale.cu:169-201,249-268)
    bool any_floor = false, any_fine = false;
    for (const Rec& e : rec_) {
      if (e.status != 0) continue;
      any_floor = any_floor || e.k >= 1;
      any_fine = any_fine || e.fine;
    }
    if (any_floor) {
      const int32_t* fon = upload_on([](const Rec& e) { return e.status == 0 && e.k >= 1; });
      int32_t mWf;
      const int32_t* dWf = upload_win([](const Rec& e) { return e.W; }, [](const Rec& e) { return e.status == 0 && e.k >= 1; }, &mWf);
      std::vector<int32_t> kk(B, 0);
      for (int64_t b = 0; b < B; ++b) kk[b] = rec_[b].k;
      const int32_t* dk = ctx_.up.put(kk);
      T* chi = X_[1];
      each(s_, g, fon, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        chi[at] = (A[at] & 1u) ? T(1) : T(0);
      });
      enqueue_cube<T>(s_, g_, CubeOp::kBox, chi, chi, nullptr, dWf, mWf);
      const T log_absent = static_cast<T>(std::log(c.absent));
      const double growth = c.growth;
      const int32_t growth_from = c.growth_from;
      const T* Lp = Lp_;
      each(s_, g, fon, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        const T lp = Lp[at];
        const bool grow = growth > 1.0 && dk[b] >= growth_from;
        T fl = grow ? static_cast<T>(growth) * lp : lp + T(2);
        const T cap = static_cast<T>(c_hi * static_cast<double>(dw[b]));
        fl = fl < cap ? fl : cap;
        const T x = chi[at];
        const T e = (T(1) - x) * log(fl) + x * log_absent;
        const T f = exp(e);
        const T l = L[at];
        L[at] = l >= f ? l : f;
      });
    }
    if (any_fine) {
      const int32_t* fon = upload_on([](const Rec& e) { return e.status == 0 && e.fine; });
      const double* dg3 = upload_f64([](const Rec& e) { return e.g3; }, [](const Rec& e) { return e.status == 0 && e.fine; });
      each(s_, g, fon, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        const T v = static_cast<T>(dg3[b]) * L[at];
        L[at] = v >= T(3) ? v : T(3);
      });
    }
  }

  // step 6: the orientation and the element build (reference driver.cpp:540-569; element.cu:166-216)
  if (D >= 2) {
    std::vector<int32_t> wv(B, 0);
    for (int64_t b = 0; b < B; ++b)
      if (rec_[b].status == 0) wv[b] = rec_[b].w;
    const int32_t* dw = ctx_.up.put(wv);
    EDA_HIP(hipMemsetAsync(d_nonfinite_, 0, sizeof(int32_t) * B, s_));
    enqueue_orientation<T>(s_, ctx_.device, g_, r_, dw, on, nvec_, coh_, owork_, d_nonfinite_, d_oscr_);
    if (orient_ != nullptr) store_orientation(on);
  }
  {
    std::vector<uint64_t> init(2 * B);
    for (int64_t b = 0; b < B; ++b) {
      init[2 * b] = 0x7ff0000000000000ULL;
      init[2 * b + 1] = 0;
    }
    const uint64_t* dinit = ctx_.up.put(init);
    flush_uploads();
    EDA_HIP(hipMemcpyAsync(d_dminmax_, dinit, sizeof(uint64_t) * 2 * B, hipMemcpyDeviceToDevice, s_));
  }
  enqueue_element_build<T>(s_, g_, c_, L_, nvec_, coh_, o_, on, d_dminmax_);
  std::vector<uint64_t> dmm(2 * B);
  std::vector<int32_t> onf(B, 0);
  d2h(dmm.data(), d_dminmax_, 2 * B);
  if (D >= 2) d2h(onf.data(), d_nonfinite_, B);
  sync();
  // step 7: the ladder, the weights, the rung windows
  std::vector<ElemDev> el(B);
  for (int64_t b = 0; b < B; ++b) {
    Rec& e = rec_[b];
    std::memset(&el[b], 0, sizeof(ElemDev));
    if (e.status != 0) continue;
    if (onf[b]) {
      e.status = 2;
      e.err = 2;
      continue;
    }
    const std::vector<int32_t> lad = rung_ladder(bits_d(dmm[2 * b]), bits_d(dmm[2 * b + 1]), e.w, c.n_max);
    el[b].n_rungs = static_cast<int32_t>(lad.size());
    for (size_t j = 0; j < lad.size(); ++j) el[b].rung_w[j] = lad[j];
  }
  any = false;
  for (const Rec& e : rec_) any = any || e.status == 0;
  if (!any) return false;
  on = upload_on(active);
  {
    const ElemDev* del = ctx_.up.put(el);
    flush_uploads();
    EDA_HIP(hipMemcpyAsync(d_el_, del, sizeof(ElemDev) * B, hipMemcpyDeviceToDevice, s_));
  }
  EDA_HIP(hipMemsetAsync(d_bits_, 0, sizeof(uint32_t) * 2 * B, s_));
  enqueue_element_weights<T>(s_, g_, c_, L_, coh_, d_el_, lo_, t_, on, d_bits_);
  std::vector<uint32_t> bits(2 * B);
  d2h(bits.data(), d_bits_, 2 * B);
  sync();
  K_max_ = 0;
  for (int64_t b = 0; b < B; ++b) {
    if (rec_[b].status != 0) continue;
    el[b].used = bits[2 * b];
    el[b].pts = bits[2 * b + 1];
    K_max_ = std::max(K_max_, el[b].n_rungs);
  }
  std::vector<int32_t> rw(static_cast<size_t>(K_max_) * B, 0);
  rung_max_.assign(static_cast<size_t>(K_max_), 0);
  for (int j = 0; j < K_max_; ++j)
    for (int64_t b = 0; b < B; ++b) {
      const ElemDev& e = el[b];
      if (rec_[b].status == 0 && j < e.n_rungs && ((e.used >> j) & 1u)) {
        rw[j * B + b] = e.rung_w[j];
        rung_max_[j] = std::max(rung_max_[j], e.rung_w[j]);
      }
    }
  {
    const ElemDev* del = ctx_.up.put(el);
    flush_uploads();
    EDA_HIP(hipMemcpyAsync(d_el_, del, sizeof(ElemDev) * B, hipMemcpyDeviceToDevice, s_));
    if (K_max_ > 0) {
      const int32_t* drw = ctx_.up.put(rw);
      flush_uploads();
      EDA_HIP(hipMemcpyAsync(d_rung_win_, drw, sizeof(int32_t) * rw.size(), hipMemcpyDeviceToDevice, s_));
    }
  }
  rw_ = rw;
  rw_cur_ = rw;
  if (!rw_cur_.empty()) mirror_note(d_rung_win_, rw_cur_.data(), rw_cur_.size() * sizeof(int32_t));
  on = upload_on(active);
  on_op_ = on;
  TapeMode* tm = nullptr;
  if (tape_ != nullptr) {
    tape_->modes.emplace_back();
    tm = &tape_->modes.back();
    tm->active.assign(B, 0);
    tm->w.assign(B, 0);
    tm->W.assign(B, 0);
    tm->tol.assign(B, 0.0);
    for (int64_t b = 0; b < B; ++b) {
      const Rec& e = rec_[b];
      if (e.status != 0) continue;
      tm->active[b] = 1;
      tm->w[b] = e.w;
      tm->W[b] = e.W;
      tm->tol[b] = e.tol;
    }
    tm->el = el;
    tm->K_max = K_max_;
    tm->rung_win = rw_;
    tm->rung_max = rung_max_;
    const size_t fb = static_cast<size_t>(B) * g.count;
    tm->r = TapeBuf(fb * sizeof(T));
    tm->L = TapeBuf(fb * sizeof(T));
    tm->lo = TapeBuf(fb);
    EDA_HIP(hipMemcpyAsync(tm->r.p, r_, fb * sizeof(T), hipMemcpyDeviceToDevice, s_));
    EDA_HIP(hipMemcpyAsync(tm->L.p, L_, fb * sizeof(T), hipMemcpyDeviceToDevice, s_));
    EDA_HIP(hipMemcpyAsync(tm->lo.p, lo_, fb, hipMemcpyDeviceToDevice, s_));
  }

  // step 8: sift 1 and the continuation weight (reference driver.cpp:606-638; sift.cu:55-85)
  T* const* X = X_;
  T* r = r_;
  T* h = h_;
  // h = r - m, and R = max(u - l, 0) for the continuation weight, out of the gathers
  envelope_mean(r_, X[0], X[1], X[2], nullptr, c.sifts > 1 ? X[5] : nullptr, h);
  for (Rec& e : rec_)
    if (e.status == 0) e.n_sifts = 1;
  T* a = X[8];
  const double* dtol = upload_f64([](const Rec& e) { return e.tol; }, active);
  if (c.sifts > 1) {
    T* R = X[5];
    T* U = R;
    T* pp[2] = {X[6], X[7]};
    for (int32_t it = 0; it < c.u; ++it) {
      T* dst = pp[it & 1];
      elem_op(0, U, dst, nullptr);
      U = dst;
    }
    T* gw = U == X[6] ? X[7] : X[6];
    const T t0 = static_cast<T>(c.t0);
    const double dd = c.t1 - c.t0;
    const T den = static_cast<T>(dd >= 1e-12 ? dd : 1e-12);
    const T* Uc = U;
    each(s_, g, on, [=] __device__(int64_t b, int64_t i) {
      const int64_t at = b * g.count + i;
      const T Uv = Uc[at];
      T rho = T(0);
      if (static_cast<double>(Uv) > dtol[b]) rho = clip_np<T>(R[at] / Uv, T(0), T(1));
      gw[at] = clip_np<T>((rho - t0) / den, T(0), T(1));
    });
    T* cur = gw;
    for (int32_t it = 0; it < c.s; ++it) {
      T* dst = a;
      if (it + 1 < c.s) dst = cur == X[5] ? X[6] : X[5];
      elem_op(3, cur, dst, nullptr);
      cur = dst;
    }
    each(s_, g, on, [=] __device__(int64_t b, int64_t i) {
      const int64_t at = b * g.count + i;
      a[at] = clip_np<T>(a[at], T(0), T(1));
    });
  }

  // step 9: sifts 2..S with G1 (reference driver.cpp:641-680; sift.cu:88-232)
  for (int32_t j = 1; j < c.sifts; ++j) {
    auto sifting = [](const Rec& e) { return e.status == 0 && e.sifting; };
    bool anys = false;
    for (const Rec& e : rec_) anys = anys || sifting(e);
    if (!anys) break;
    const int32_t* son = upload_on(sifting);
    on_op_ = son;
    if (j > 1) {
      // elements that stopped sifting drop out of the operators' cubes and gathers
      std::vector<int32_t> rw = rw_;
      for (int jj = 0; jj < K_max_; ++jj)
        for (int64_t b = 0; b < B; ++b)
          if (!sifting(rec_[b])) rw[jj * B + b] = 0;
      const int32_t* drw = ctx_.up.put(rw);
      flush_uploads();
      EDA_HIP(hipMemcpyAsync(d_rung_win_, drw, sizeof(int32_t) * rw.size(), hipMemcpyDeviceToDevice, s_));
      rw_cur_ = rw;
      if (!rw_cur_.empty()) mirror_note(d_rung_win_, rw_cur_.data(), rw_cur_.size() * sizeof(int32_t));
    }
    T* plain = X[0];
    T* up = X[1];
    T* lo = X[2];
    T* hd = X[5];
    envelope_mean(h, plain, up, lo, hd);
    T* e0 = X[6];
    elem_op(3, hd, e0, nullptr);
    int32_t mWs;
    const int32_t* dWs = upload_win([](const Rec& e) { return e.W; }, sifting, &mWs);
    const double* dtol = upload_f64([](const Rec& e) { return e.tol; }, sifting);
    T* emax = X[5];
    T* emin = X[7];
    enqueue_cube<T>(s_, g_, CubeOp::kMaxMin, e0, emax, emin, dWs, mWs);
    {
      const T k0 = static_cast<T>(c.ramp_k0);
      const double dd = c.ramp_k1 - c.ramp_k0;
      const T kden = static_cast<T>(dd >= 1e-12 ? dd : 1e-12);
      const double tiny = c.tiny;
      each(s_, g, son, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        const double tl = dtol[b];
        const T fl = static_cast<T>(tl >= tiny ? tl : tiny);
        const T hi = emax[at], lw = emin[at];
        const T sum = hi + lw;
        const T den = sum >= fl ? sum : fl;
        const T kap = (hi - lw) / den;
        emax[at] = clip_np<T>((kap - k0) / kden, T(0), T(1));
      });
    }
    T* gk = X[7];
    enqueue_cube<T>(s_, g_, CubeOp::kBox, emax, gk, nullptr, dWs, mWs);
    each(s_, g, son, [=] __device__(int64_t b, int64_t i) {
      const int64_t at = b * g.count + i;
      gk[at] = clip_np<T>(gk[at], T(0), T(1));
    });
    T* e = X[9];
    if (c.anorm_smooth <= 0) {
      EDA_HIP(hipMemcpyAsync(e, e0, sizeof(T) * B * g.count, hipMemcpyDeviceToDevice, s_));
    }
    for (int32_t it = 0; it < c.anorm_smooth; ++it) enqueue_cube<T>(s_, g_, CubeOp::kBox, it ? e : e0, e, nullptr, dWs, mWs);
    T* q = X[6];
    {
      Selection<T> sel{};
      sel.src = SelSrc::kField;
      sel.field = e;
      enqueue_select<T>(s_, g_, sel, 0, 0.5, d_e_med_, nullptr, son, sel_ws_);
      const double* emed = d_e_med_;
      const double floor_k = c.anorm_floor, tiny = c.tiny;
      each(s_, g, son, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        double thr = floor_k * emed[b];
        const double tl = dtol[b];
        thr = thr >= tl ? thr : tl;
        thr = thr >= tiny ? thr : tiny;
        const T ev = e[at];
        const T ef = static_cast<double>(ev) >= thr ? ev : static_cast<T>(thr);
        e[at] = ef;
        q[at] = h[at] / ef;
      });
    }
    T* Mn = X[5];
    envelope_mean(q, Mn, X[10], X[11]);
    each(s_, g, son, [=] __device__(int64_t b, int64_t i) {
      const int64_t at = b * g.count + i;
      const T wv = gk[at];
      const T mA = clip_np<T>(e[at] * Mn[at], lo[at], up[at]);
      Mn[at] = wv * mA + (T(1) - wv) * plain[at];
    });
    // the sift with the sup-norm guard: hn = h - a mh; accepted iff max|r - hn| <= r_inf and
    // max|hn| <= r_inf (reference sift.cu:160-232)
    T* hn = X[6];
    uint64_t* gmax = d_gmax_;
    EDA_HIP(hipMemsetAsync(gmax, 0, sizeof(uint64_t) * 2 * B, s_));
    {
      int64_t gx = (g.count + kThreads - 1) / kThreads;
      gx = gx > 64 ? 64 : gx;  // 64 workgroups per element: 64 atomics per element and maximum
      EDA_GO(k_sift_step<T>, dim3(static_cast<unsigned>(gx), static_cast<unsigned>(B)), dim3(kThreads), 0,
                         s_, g.count, son, a, Mn, h, r, hn, gmax);
      EDA_LAUNCH("k_sift_step");
    }
    const double* drinf = upload_f64([](const Rec& e) { return e.r_inf; }, sifting);
    each(s_, g, son, [=] __device__(int64_t b, int64_t i) {
      const bool ok = !(bits_d(gmax[2 * b]) > drinf[b]) && !(bits_d(gmax[2 * b + 1]) > drinf[b]);
      if (ok) {
        const int64_t at = b * g.count + i;
        h[at] = hn[at];
      }
    });
    std::vector<uint64_t> gm(2 * B);
    d2h(gm.data(), gmax, 2 * B);
    std::vector<double> emed_h;
    if (tm != nullptr) {
      emed_h.resize(B);
      d2h(emed_h.data(), d_e_med_, B);
    }
    sync();
    if (tm != nullptr) {
      tm->ran.emplace_back(B, 0);
      tm->acc.emplace_back(B, 0);
      tm->thr.emplace_back(B, 0.0);
    }
    for (int64_t b = 0; b < B; ++b) {
      Rec& er = rec_[b];
      if (!sifting(er)) continue;
      const bool ok = !(bits_d(gm[2 * b]) > er.r_inf) && !(bits_d(gm[2 * b + 1]) > er.r_inf);
      if (tm != nullptr) {
        // the floor as the device computed it: max(anorm_floor median(e), tol, tiny) in double
        double thr = c.anorm_floor * emed_h[b];
        thr = thr >= er.tol ? thr : er.tol;
        thr = thr >= c.tiny ? thr : c.tiny;
        tm->ran.back()[b] = 1;
        tm->acc.back()[b] = ok ? 1 : 0;
        tm->thr.back()[b] = thr;
      }
      if (ok) ++er.n_sifts;
      else er.sifting = false;
    }
  }

  // step 10: the commit (reference mode.cu:306-348,420-430; field_ops.cu:103-154)
  {
    on = upload_on(active);
    std::vector<int32_t> row(B, 0);
    for (int64_t b = 0; b < B; ++b) row[b] = rec_[b].n_imfs;
    const int32_t* drow = ctx_.up.put(row);
    T* rows = rows_;
    const int64_t cap = cap_;
    each(s_, g, on, [=] __device__(int64_t b, int64_t i) {
      const int64_t at = b * g.count + i;
      const T hv = h[at];
      rows[(b * cap + drow[b]) * g.count + i] = hv;
      r[at] = r[at] - hv;
    });
    enqueue_elem_stats<T>(s_, g_, h_, on, d_energy_, nullptr, nullptr, stats_ws_);
    std::vector<double> en(B);
    d2h(en.data(), d_energy_, B);
    enqueue_elem_stats<T>(s_, g_, r_, on, nullptr, d_amax_, nullptr, stats_ws_);
    std::vector<uint64_t> am(B);
    d2h(am.data(), d_amax_, B);
    sync();
    for (int64_t b = 0; b < B; ++b) {
      Rec& e = rec_[b];
      if (e.status != 0) continue;
      const int32_t k = e.n_imfs;
      hist_w_[b][k] = e.w;
      hist_e_[b][k] = en[b];
      hist_sp_[b][k] = e.spacing;
      hist_s_[b][k] = e.n_sifts;
      e.n_imfs = k + 1;
      e.amax = bits_d(am[b]);
      e.tol = tie_tol<T>(tau_, e.amax);
      if (max_imf_ > 0 && e.n_imfs >= max_imf_) e.status = 1;
      if (e.n_imfs + 1 >= cap_) throw std::logic_error("eda: the output rows are full (mode bound violated)");
      e.w_prev = e.w;
      e.k += 1;
      if (!std::isfinite(e.amax)) {
        e.status = 2;
        e.err = 2;
      }
    }
  }
  std::swap(L_, Lp_);
  any = false;
  for (const Rec& e : rec_) any = any || e.status == 0;
  return any;
}

// ------------------------------------------------------------------------------------------------
// Sets the calling thread's DCT contraction mode for a scope.
struct DctModeScope {
  int prev;
  explicit DctModeScope(int mode) : prev(set_dct_mode(mode)) {}
  ~DctModeScope() { set_dct_mode(prev); }
};

template <class T>
void run_call(const eda_array_desc& x, const eda_params& p, Run& run) {
  Context& ctx = context(run.device);
  if (p.dct_mode != 0 && p.dct_mode != 1) throw std::invalid_argument("eda: dct_mode must be 0 or 1");
  const DctModeScope dct_scope(std::is_same_v<T, float> ? p.dct_mode : 0);
  // device input is read on the library's stream (non-blocking) after the caller's stream: handle 0
  // is the null stream, which a non-blocking stream does not wait for by itself
  if (p.stream || x.device >= 0) {
    hipEvent_t ev;
    EDA_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    EDA_HIP(hipEventRecord(ev, reinterpret_cast<hipStream_t>(p.stream)));
    EDA_HIP(hipStreamWaitEvent(ctx.stream, ev, 0));
    EDA_HIP(hipEventDestroy(ev));
  }
  const Geom gall = make_geom(x.dim, x.batch, x.n);
  MethodCtl c = method_ctl(gall, std::is_same_v<T, float>);
  c.sifts = p.sifts;
  const int32_t bound = mode_bound(c);
  const int32_t modes_max = p.max_imf > 0 ? std::min(bound, p.max_imf) : bound;
  run.cap = modes_max + 2;
  const double tau = p.tie_tolerance >= 0 ? p.tie_tolerance : (std::is_same_v<T, float> ? std::ldexp(1.0, -14) : std::ldexp(1.0, -30));
  const int64_t B = x.batch;
  run.rows = DevMem(static_cast<size_t>(B) * run.cap * gall.count * sizeof(T), ctx.stream);
  const size_t he = static_cast<size_t>(B) * (run.cap - 1);
  run.windows.assign(he, 0);
  run.sifts.assign(he, 0);
  run.spacing.assign(he, 0.0);
  run.energy.assign(he, 0.0);
  run.n_imfs.assign(B, 0);
  run.error.assign(B, 0);
  if (p.orientation && x.dim >= 2) {
    run.orient = DevMem(static_cast<size_t>(B) * (run.cap - 1) * (x.dim + 1) * gall.count * sizeof(T), ctx.stream);
    run.kept_orient = true;
  }
  if (p.keep_tape) {
    run.tape = std::make_unique<Tape>();
    run.tape->dtype = std::is_same_v<T, float> ? EDA_F32 : EDA_F64;
    run.tape->dct_mode = std::is_same_v<T, float> ? p.dct_mode : 0;
    run.tape->c = c;
  }
  // the chunk: bounded by max_chunk and by the free device memory (halving)
  int64_t Bc = std::min<int64_t>(B, 65535);  // the kernels put the elements of a chunk on grid.y
  if (p.max_chunk > 0) Bc = std::min<int64_t>(Bc, p.max_chunk);
  size_t free_b = 0, total_b = 0;
  EDA_HIP(hipMemGetInfo(&free_b, &total_b));
  free_b += BlockCache::get().held(run.device);  // cached free blocks serve the call too
  const size_t per_elem = static_cast<size_t>(gall.count) * (sizeof(T) * 40 + 16) + 4096;
  while (Bc > 1 && static_cast<size_t>(Bc) * per_elem + (size_t{512} << 20) > free_b) Bc = (Bc + 1) / 2;
  run.chunk = Bc;
  for (int64_t b0 = 0; b0 < B; b0 += Bc) {
    const int64_t nb = std::min(Bc, B - b0);
    const Geom g = make_geom(x.dim, nb, x.n);
    T* rows = static_cast<T*>(run.rows.p) + static_cast<size_t>(b0) * run.cap * g.count;
    T* orient = run.kept_orient ? static_cast<T*>(run.orient.p) +
                                      static_cast<size_t>(b0) * (run.cap - 1) * (x.dim + 1) * g.count
                                : nullptr;
    Chunk<T> ch(ctx, g, c, tau, p.check_finite != 0, p.max_imf, rows, run.cap, orient);
    ChunkTape* ct = nullptr;
    if (run.tape) {
      run.tape->chunks.emplace_back();
      ct = &run.tape->chunks.back();
      ct->b0 = b0;
      ct->g = g;
      ch.keep_tape(ct);
    }
    ch.alloc_history();
    ch.start(x, b0);
    while (ch.mode()) {
    }
    ch.finish(run, b0);
    if (ct != nullptr) ch.tape_results(*ct, run, b0);
  }
  EDA_HIP(hipStreamSynchronize(ctx.stream));
  int32_t P = 0, K = 0;
  bool any_err = false;
  for (int64_t b = 0; b < B; ++b) {
    if (run.error[b]) any_err = true;
    else P = std::max(P, run.n_imfs[b] + 1);
    K = std::max(K, run.n_imfs[b]);
  }
  if (any_err) P = std::max(P, 1);
  run.P = P;
  run.K = K;
}

// ------------------------------------------------------------------------------------------------
// The C ABI plumbing.
void set_err(char* err, size_t len, const std::string& msg) {
  if (err == nullptr || len == 0) return;
  std::snprintf(err, len, "%s", msg.c_str());
}
template <class F>
int guarded(char* err, size_t len, F&& f) {
  try {
    f();
    return EDA_OK;
  } catch (const OutOfDeviceMemory& e) {
    set_err(err, len, e.what());
    return EDA_ENOMEM;
  } catch (const std::bad_alloc&) {
    set_err(err, len, "eda: host memory exhausted");
    return EDA_ENOMEM;
  } catch (const HipError& e) {
    set_err(err, len, e.what());
    return EDA_EHIP;
  } catch (const std::invalid_argument& e) {
    set_err(err, len, e.what());
    return EDA_EINVAL;
  } catch (const std::exception& e) {
    set_err(err, len, e.what());
    return EDA_EINTERNAL;
  }
}

}  // namespace

int abi_guard(char* err, size_t errlen, const std::function<void()>& f) { return guarded(err, errlen, f); }

namespace {

void validate(const eda_array_desc& x, const eda_params& p) {
  if (x.dim < 1 || x.dim > 3) throw std::invalid_argument("eda: D must be 1, 2 or 3");
  if (x.dtype != EDA_F32 && x.dtype != EDA_F64) throw std::invalid_argument("eda: the input must be float32 or float64");
  // 1-D lines take the long-line paths of the cube filter, the extrema and the reductions (rows of
  // up to 2^18 samples); fields of 2 or 3 axes keep the limits of their staged kernels and transforms
  const int64_t lim = x.dim == 1 ? (int64_t{1} << 18) : (x.dtype == EDA_F32 ? 8192 : 4096);
  for (int a = 0; a < x.dim; ++a) {
    if (x.n[a] < 1) throw std::invalid_argument("eda: every spatial extent must be >= 1");
    if (x.n[a] > lim) {
      throw std::invalid_argument("eda: axis extent " + std::to_string(x.n[a]) + " exceeds the limit " +
                                  std::to_string(lim) + " for " + (x.dtype == EDA_F32 ? "float32" : "float64"));
    }
  }
  for (int a = 0; a < x.dim; ++a) {
    if (x.n[a] < 2) {
      if (x.dim == 1) {
        throw std::invalid_argument(
            "eda: a 1D signal needs at least 2 samples, got 1 (one sample has no sample differences, so no noise "
            "level; it is its own residue)");
      }
      throw std::invalid_argument("eda: every spatial axis of a " + std::to_string(x.dim) +
                                  "D field needs extent >= 2, but spatial axis " + std::to_string(a) +
                                  " has extent 1; drop that axis (squeeze it, or leave it out of axes=) to decompose "
                                  "the field in " + std::to_string(x.dim - 1) + "D");
    }
  }
  int64_t count = 1;
  for (int a = 0; a < x.dim; ++a) count *= x.n[a];
  if (count > (int64_t{1} << 31)) throw std::invalid_argument("eda: an element has more than 2^31 samples");
  if (x.batch < 0) throw std::invalid_argument("eda: the batch must be >= 0");
  if (p.sifts < 1) throw std::invalid_argument("eda: sifts must be >= 1");
  if (p.max_chunk < 0) throw std::invalid_argument("eda: max_chunk must be >= 0");
  if (p.tie_tolerance >= 0 && !std::isfinite(p.tie_tolerance))
    throw std::invalid_argument("eda: tie_tolerance must be a finite value >= 0");
  if (x.batch > 0 && x.data == nullptr) throw std::invalid_argument("eda: the input has no data pointer");
  if (x.device < 0 && x.source != EDA_SRC_SAME)
    throw std::invalid_argument("eda: host input must already be in the compute dtype");
  const size_t item = source_size(x.source, x.dtype);
  if (x.device >= 0 && x.batch > 0 && reinterpret_cast<uintptr_t>(x.data) % item != 0) {
    throw std::invalid_argument("eda: the device array's data pointer is not aligned to its item size (" +
                                std::to_string(item) + " bytes)");
  }
  if (x.device >= 0 && x.batch_axes > 1) {
    if (x.batch_axes > 64) throw std::invalid_argument("eda: a device array may have at most 64 batch axes");
    if (x.batch_shape == nullptr || x.batch_strides == nullptr)
      throw std::invalid_argument("eda: batch_axes > 1 needs batch_shape and batch_strides");
    int64_t prod = 1;
    for (int k = 0; k < x.batch_axes; ++k) {
      if (x.batch_shape[k] < 0) throw std::invalid_argument("eda: negative extent");
      prod *= x.batch_shape[k];
    }
    if (prod != x.batch) throw std::invalid_argument("eda: the batch shape does not hold `batch` elements");
  }
}

template <class T>
void collect_rows(const Run& run, void* dst, int32_t dst_device) {
  Context& ctx = context(run.device);
  const int64_t B = run.batch, count = run.count;
  const int32_t P = run.P, cap = run.cap;
  if (B == 0 || P == 0) return;
  const size_t elems = static_cast<size_t>(B) * P * count;
  DevMem tmp;
  T* out;
  if (dst_device >= 0) {
    out = static_cast<T*>(dst);
  } else {
    tmp = DevMem(elems * sizeof(T), ctx.stream);
    out = static_cast<T*>(tmp.p);
  }
  const int32_t* dn = ctx.up.put(run.n_imfs);
  const int32_t* de = ctx.up.put(run.error);
  const T* src = static_cast<const T*>(run.rows.p);
  // row k of element b (bk = b P + k) over a (B P, count) grid
  Geom gr{};
  gr.count = count;
  gr.batch = B * P;
  each(ctx.stream, gr, nullptr, [=] __device__(int64_t bk, int64_t i) {
    const int64_t b = bk / P;
    const int64_t k = bk - b * P;
    T v;
    if (de[b] != 0) v = static_cast<T>(__longlong_as_double(0x7ff8000000000000LL));
    else if (k <= dn[b]) v = src[(b * cap + k) * count + i];
    else v = T(0);
    out[bk * count + i] = v;
  });
  if (dst_device < 0) EDA_HIP(hipMemcpyAsync(dst, out, elems * sizeof(T), hipMemcpyDeviceToHost, ctx.stream));
  EDA_HIP(hipStreamSynchronize(ctx.stream));
  ctx.up.reset();
}

// dir (B, K, D, count) and coh (B, K, count) <- rows 0 .. n_imfs[b] - 1 of the kept orientation; 0
// above them and for failed elements (reference orientation.cu:524-549, driver.cpp:177-196).
template <class T>
void collect_orientation(const Run& run, void* dir, void* coh, int32_t dst_device, int32_t K) {
  Context& ctx = context(run.device);
  const int64_t B = run.batch, count = run.count;
  const int D = run.dim;
  const int64_t rows = run.cap - 1;
  const size_t nd = static_cast<size_t>(B) * K * D * count, nc = static_cast<size_t>(B) * K * count;
  DevMem tmp;
  T *od, *oc;
  if (dst_device >= 0) {
    od = static_cast<T*>(dir);
    oc = static_cast<T*>(coh);
  } else {
    tmp = DevMem((nd + nc) * sizeof(T), ctx.stream);
    od = static_cast<T*>(tmp.p);
    oc = od + nd;
  }
  const int32_t* dn = ctx.up.put(run.n_imfs);
  const int32_t* de = ctx.up.put(run.error);
  const T* src = static_cast<const T*>(run.orient.p);
  flat(ctx.stream, static_cast<int64_t>(nd + nc), [=] __device__(int64_t idx) {
    const int64_t i = idx % count;
    const int64_t q = idx / count;
    const int64_t a = q % (D + 1);
    const int64_t bk = q / (D + 1);
    const int64_t k = bk % K;
    const int64_t b = bk / K;
    const T v = de[b] == 0 && k < dn[b] ? src[((b * rows + k) * (D + 1) + a) * count + i] : T(0);
    if (a < D) od[(bk * D + a) * count + i] = v;
    else oc[bk * count + i] = v;
  });
  if (dst_device < 0) {
    EDA_HIP(hipMemcpyAsync(dir, od, nd * sizeof(T), hipMemcpyDeviceToHost, ctx.stream));
    EDA_HIP(hipMemcpyAsync(coh, oc, nc * sizeof(T), hipMemcpyDeviceToHost, ctx.stream));
  }
  EDA_HIP(hipStreamSynchronize(ctx.stream));
  ctx.up.reset();
}

int current_device_of(const eda_array_desc& x);

// eda.monogenic (reference cpp/src/method/monogenic.cpp:61-158): per chunk of elements sized by
// the free device memory, the input copy (host staged or device gathered through its strides), the
// non-finite counts and max |x|, the exact prescale f 2^-k with max |f| in [1, 2), and the monogenic
// signal into the call's planes. Returns the chunk.
template <class T>
int64_t run_monogenic(const eda_array_desc& x, bool check_finite, int64_t max_chunk, uint64_t stream, void* out,
                      int32_t out_device, int32_t* nonfinite) {
  const int device = current_device_of(x);
  Context& ctx = context(device);
  hipStream_t s = ctx.stream;
  if (stream || x.device >= 0) {  // as in run_call: 0 is the null stream
    hipEvent_t ev;
    EDA_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    EDA_HIP(hipEventRecord(ev, reinterpret_cast<hipStream_t>(stream)));
    EDA_HIP(hipStreamWaitEvent(s, ev, 0));
    EDA_HIP(hipEventDestroy(ev));
  }
  const Geom gall = make_geom(x.dim, x.batch, x.n);
  const int D = x.dim;
  const int64_t B = x.batch, count = gall.count;
  const size_t out_elems = static_cast<size_t>(B) * monogenic_planes(D) * count;
  DevMem dout;
  T* o = static_cast<T*>(out);
  if (out_device < 0) {
    dout = DevMem(out_elems * sizeof(T), s);
    o = static_cast<T*>(dout.p);
  }
  const double tiny = method_ctl(gall, std::is_same_v<T, float>).tiny;
  int64_t Bc = std::min<int64_t>(B, 65535);
  if (max_chunk > 0) Bc = std::min(Bc, max_chunk);
  const auto ws_bytes = [&](int64_t nb) {
    const Geom g = make_geom(x.dim, nb, x.n);
    const size_t F = static_cast<size_t>(nb) * count;
    return (monogenic_fields(D) + 1) * (F * sizeof(T) + 256) + elem_stats_ws_bytes(g) + 256 +
           static_cast<size_t>(orient_scratch_doubles(nb)) * sizeof(double) + static_cast<size_t>(nb) * 24 + 4096;
  };
  size_t free_b = 0, total_b = 0;
  EDA_HIP(hipMemGetInfo(&free_b, &total_b));
  free_b += BlockCache::get().held(device);
  while (Bc > 1 && ws_bytes(Bc) + (size_t{512} << 20) > free_b) Bc = (Bc + 1) / 2;
  DevMem mem(ws_bytes(Bc), s);
  std::vector<uint64_t> amax(Bc), nf(Bc);
  std::vector<int32_t> kexp(Bc);
  for (int64_t b0 = 0; b0 < B; b0 += Bc) {
    const int64_t nb = std::min(Bc, B - b0);
    const Geom g = make_geom(x.dim, nb, x.n);
    const int64_t F = nb * count;
    Carve cv{static_cast<char*>(mem.p)};
    T* f = cv.take<T>(F);
    T* work = cv.take<T>(static_cast<size_t>(F) * monogenic_fields(D));
    char* sws = cv.take<char>(elem_stats_ws_bytes(g));
    double* dscr = cv.take<double>(static_cast<size_t>(orient_scratch_doubles(nb)));
    uint64_t* d_amax = cv.take<uint64_t>(nb);
    uint64_t* d_nf = cv.take<uint64_t>(nb);
    ctx.up.reset();
    fill_input<T>(ctx, g, x, b0, f);
    enqueue_elem_stats<T>(s, g, f, nullptr, nullptr, d_amax, d_nf, sws);
    EDA_HIP(hipMemcpyAsync(amax.data(), d_amax, sizeof(uint64_t) * nb, hipMemcpyDeviceToHost, s));
    EDA_HIP(hipMemcpyAsync(nf.data(), d_nf, sizeof(uint64_t) * nb, hipMemcpyDeviceToHost, s));
    EDA_HIP(hipStreamSynchronize(s));
    for (int64_t b = 0; b < nb; ++b) {
      const double a = bits_d(amax[b]);
      kexp[b] = (a > 0.0 && std::isfinite(a)) ? std::ilogb(a) : 0;
      if (nonfinite != nullptr) nonfinite[b0 + b] = check_finite && nf[b] != 0 ? 1 : 0;
    }
    const int32_t* dk = ctx.up.put(std::vector<int32_t>(kexp.begin(), kexp.begin() + nb));
    each(s, g, nullptr, [=] __device__(int64_t b, int64_t i) {
      const int32_t k = dk[b];
      if (k == 0) return;
      const int64_t at = b * g.count + i;
      if constexpr (std::is_same_v<T, float>) f[at] = static_cast<float>(ldexp(static_cast<double>(f[at]), -k));
      else f[at] = ldexp(f[at], -k);
    });
    enqueue_monogenic<T>(s, device, g, f, dk, tiny, work, dscr, o, B, b0);
    EDA_HIP(hipStreamSynchronize(s));
  }
  if (out_device < 0 && out_elems)
    EDA_HIP(hipMemcpyAsync(out, o, out_elems * sizeof(T), hipMemcpyDeviceToHost, s));
  EDA_HIP(hipStreamSynchronize(s));
  ctx.up.reset();
  return Bc;
}

int current_device_of(const eda_array_desc& x) {
  if (x.device >= 0) return x.device;
  int d = 0;
  EDA_HIP(hipGetDevice(&d));
  return d;
}

template <class T>
struct PrimGeom {
  Geom g;
  PrimGeom(int32_t dim, int64_t batch, const int64_t* n) {
    if (dim < 1 || dim > 3 || batch < 1) throw std::invalid_argument("eda: prim: bad geometry");
    g = make_geom(dim, batch, n);
  }
};

}  // namespace
}  // namespace eda

// ================================================================================================
using namespace eda;

extern "C" {

int eda_decompose(const eda_array_desc* x, const eda_params* p, eda_result* out, char* err, size_t errlen) {
  return guarded(err, errlen, [&] {
    if (x == nullptr || p == nullptr || out == nullptr) throw std::invalid_argument("eda: null argument");
    std::memset(out, 0, sizeof *out);
    validate(*x, *p);
    auto run = std::make_unique<Run>();
    run->device = current_device_of(*x);
    run->dtype = x->dtype;
    run->dim = x->dim;
    run->batch = x->batch;
    run->count = 1;
    for (int a = 0; a < x->dim; ++a) {
      run->n[a] = x->n[a];
      run->count *= x->n[a];
    }
    if (x->batch > 0) {
      EDA_HIP(hipSetDevice(run->device));
      if (x->dtype == EDA_F32) run_call<float>(*x, *p, *run);
      else run_call<double>(*x, *p, *run);
    } else {
      run->P = 1;
    }
    out->batch = run->batch;
    out->count = run->count;
    out->rows = run->P;
    out->K = run->K;
    out->dtype = run->dtype;
    out->device = run->device;
    out->chunk = run->chunk;
    const int64_t B = run->batch;
    const int32_t K = run->K;
    const int32_t hc = std::max(run->cap - 1, 0);
    out->n_imfs = static_cast<int32_t*>(std::calloc(std::max<int64_t>(B, 1), sizeof(int32_t)));
    out->error = static_cast<int32_t*>(std::calloc(std::max<int64_t>(B, 1), sizeof(int32_t)));
    const size_t nk = static_cast<size_t>(std::max<int64_t>(B * K, 1));
    out->windows = static_cast<int32_t*>(std::calloc(nk, sizeof(int32_t)));
    out->sifts = static_cast<int32_t*>(std::calloc(nk, sizeof(int32_t)));
    out->spacing = static_cast<double*>(std::calloc(nk, sizeof(double)));
    out->energy = static_cast<double*>(std::calloc(nk, sizeof(double)));
    if (!out->n_imfs || !out->error || !out->windows || !out->sifts || !out->spacing || !out->energy) throw std::bad_alloc();
    for (int64_t b = 0; b < B; ++b) {
      out->n_imfs[b] = run->n_imfs[b];
      out->error[b] = run->error[b];
      for (int32_t k = 0; k < K && k < hc; ++k) {
        out->windows[b * K + k] = run->windows[b * hc + k];
        out->sifts[b * K + k] = run->sifts[b * hc + k];
        out->spacing[b * K + k] = run->spacing[b * hc + k];
        out->energy[b * K + k] = run->energy[b * hc + k];
      }
    }
    out->handle = run.release();
  });
}

int eda_collect(const eda_result* r, void* dst, int32_t dst_device, char* err, size_t errlen) {
  return guarded(err, errlen, [&] {
    if (r == nullptr || r->handle == nullptr) throw std::invalid_argument("eda: collect: no result");
    const Run& run = *static_cast<const Run*>(r->handle);
    if (run.batch == 0) return;
    if (dst == nullptr) throw std::invalid_argument("eda: collect: null destination");
    if (run.dtype == EDA_F32) collect_rows<float>(run, dst, dst_device);
    else collect_rows<double>(run, dst, dst_device);
  });
}

int eda_collect_orientation(const eda_result* r, void* dir, void* coh, int32_t dst_device, char* err,
                            size_t errlen) {
  return guarded(err, errlen, [&] {
    if (r == nullptr || r->handle == nullptr) throw std::invalid_argument("eda: collect: no result");
    const Run& run = *static_cast<const Run*>(r->handle);
    if (run.batch > 0 && !run.kept_orient) throw std::logic_error("eda: the call kept no orientation");
    if (run.batch == 0 || run.K == 0) return;
    if (dir == nullptr || coh == nullptr) throw std::invalid_argument("eda: collect: null destination");
    if (run.dtype == EDA_F32) collect_orientation<float>(run, dir, coh, dst_device, run.K);
    else collect_orientation<double>(run, dir, coh, dst_device, run.K);
  });
}

int eda_monogenic(const eda_array_desc* x, int32_t check_finite, int64_t max_chunk, uint64_t stream, void* out,
                  int32_t out_device, int32_t* nonfinite, int64_t* chunk, char* err, size_t errlen) {
  return guarded(err, errlen, [&] {
    if (x == nullptr) throw std::invalid_argument("eda: null argument");
    eda_params p{};
    p.sifts = 3;
    p.tie_tolerance = -1.0;
    validate(*x, p);
    if (max_chunk < 0) throw std::invalid_argument("eda: max_chunk must be >= 0");
    if (chunk != nullptr) *chunk = 0;
    if (x->batch == 0) return;
    if (out == nullptr) throw std::invalid_argument("eda: monogenic: null output");
    EDA_HIP(hipSetDevice(current_device_of(*x)));
    const int64_t c = x->dtype == EDA_F32
                          ? run_monogenic<float>(*x, check_finite != 0, max_chunk, stream, out, out_device, nonfinite)
                          : run_monogenic<double>(*x, check_finite != 0, max_chunk, stream, out, out_device, nonfinite);
    if (chunk != nullptr) *chunk = c;
  });
}

void eda_release(eda_result* r) {
  if (r == nullptr) return;
  delete static_cast<Run*>(r->handle);
  std::free(r->n_imfs);
  std::free(r->error);
  std::free(r->windows);
  std::free(r->sifts);
  std::free(r->spacing);
  std::free(r->energy);
  std::memset(r, 0, sizeof *r);
}

int eda_build_info(char* buf, size_t len) {
  std::snprintf(buf, len,
                "{\"version\": \"0.1.0\", \"backend\": \"hip\", \"architectures\": [\"gfx950\"], "
                "\"hip_version\": %d, \"compiler\": \"%s\"}",
                HIP_VERSION, __VERSION__);
  return EDA_OK;
}

int eda_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int eda_device_info(int32_t device, char* buf, size_t len) {
  return guarded(buf, len, [&] {
    hipDeviceProp_t pr;
    EDA_HIP(hipGetDeviceProperties(&pr, device));
    std::snprintf(buf, len,
                  "{\"name\": \"%s\", \"gcn_arch\": \"%s\", \"multiprocessors\": %d, \"total_memory\": %zu, "
                  "\"lds_per_block\": %zu, \"warp_size\": %d}",
                  pr.name, pr.gcnArchName, pr.multiProcessorCount, pr.totalGlobalMem, pr.sharedMemPerBlock, pr.warpSize);
  });
}

int eda_pointer_device(const void* ptr, int32_t* device) {
  return guarded(nullptr, 0, [&] {
    hipPointerAttribute_t attr;
    const hipError_t e = hipPointerGetAttributes(&attr, ptr);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      *device = -1;
      return;
    }
    *device = attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ? attr.device : -1;
  });
}

int eda_timer_enable(int32_t on) {
  if (on < 0 || on > 15) return EDA_EINVAL;
  timer_set(on);
  return EDA_OK;
}

int eda_timer_read(double* ms, int64_t* launches, double* bytes, char* err, size_t errlen) {
  return guarded(err, errlen, [&] { timer_read(ms, launches, bytes); });
}

int eda_backward(const eda_result* r, const void* grad_rows, void* grad_x, int32_t flags, char* err, size_t errlen) {
  return guarded(err, errlen, [&] {
    if (r == nullptr || r->handle == nullptr) throw std::invalid_argument("eda: backward: no result");
    Run& run = *static_cast<Run*>(r->handle);
    if (!run.tape) throw std::invalid_argument("eda: backward: the call kept no tape (eda_params.keep_tape)");
    if (run.batch == 0) return;
    if (grad_rows == nullptr || grad_x == nullptr) throw std::invalid_argument("eda: backward: null pointer");
    Context& ctx = context(run.device);
    const DctModeScope dct_scope(run.tape->dct_mode);
    const bool orient = (flags & 1) == 0;
    for (const ChunkTape& ct : run.tape->chunks) {
      const size_t item = run.dtype == EDA_F32 ? 4 : 8;
      const char* gr = static_cast<const char*>(grad_rows) + static_cast<size_t>(ct.b0) * run.P * run.count * item;
      char* gx = static_cast<char*>(grad_x) + static_cast<size_t>(ct.b0) * run.count * item;
      if (run.dtype == EDA_F32)
        adjoint_chunk<float>(ctx.stream, run.device, *run.tape, ct, reinterpret_cast<const float*>(gr), run.P,
                             reinterpret_cast<float*>(gx), orient);
      else
        adjoint_chunk<double>(ctx.stream, run.device, *run.tape, ct, reinterpret_cast<const double*>(gr), run.P,
                              reinterpret_cast<double*>(gx), orient);
    }
  });
}

int eda_replay(const eda_result* r, const void* x, void* rows, char* err, size_t errlen) {
  return guarded(err, errlen, [&] {
    if (r == nullptr || r->handle == nullptr) throw std::invalid_argument("eda: replay: no result");
    Run& run = *static_cast<Run*>(r->handle);
    if (!run.tape) throw std::invalid_argument("eda: replay: the call kept no tape (eda_params.keep_tape)");
    if (run.batch == 0) return;
    if (x == nullptr || rows == nullptr) throw std::invalid_argument("eda: replay: null pointer");
    Context& ctx = context(run.device);
    const DctModeScope dct_scope(run.tape->dct_mode);
    for (const ChunkTape& ct : run.tape->chunks) {
      const size_t item = run.dtype == EDA_F32 ? 4 : 8;
      const char* xs = static_cast<const char*>(x) + static_cast<size_t>(ct.b0) * run.count * item;
      char* rs = static_cast<char*>(rows) + static_cast<size_t>(ct.b0) * run.P * run.count * item;
      if (run.dtype == EDA_F32)
        replay_chunk<float>(ctx.stream, run.device, *run.tape, ct, reinterpret_cast<const float*>(xs),
                            reinterpret_cast<float*>(rs), run.P);
      else
        replay_chunk<double>(ctx.stream, run.device, *run.tape, ct, reinterpret_cast<const double*>(xs),
                             reinterpret_cast<double*>(rs), run.P);
    }
  });
}

int eda_tape_mode(const eda_result* r, int64_t element, int32_t mode, int32_t* ints, double* dbls, void* residual,
                  void* L, uint8_t* lo, char* err, size_t errlen) {
  return guarded(err, errlen, [&] {
    if (r == nullptr || r->handle == nullptr) throw std::invalid_argument("eda: tape: no result");
    Run& run = *static_cast<Run*>(r->handle);
    if (!run.tape) throw std::invalid_argument("eda: tape: the call kept no tape (eda_params.keep_tape)");
    if (element < 0 || element >= run.batch) throw std::invalid_argument("eda: tape: element out of range");
    if (ints == nullptr || dbls == nullptr) throw std::invalid_argument("eda: tape: null pointer");
    for (const ChunkTape& ct : run.tape->chunks) {
      const int64_t B = static_cast<int64_t>(ct.n_imfs.size());
      if (element < ct.b0 || element >= ct.b0 + B) continue;
      const int64_t b = element - ct.b0;
      if (mode < 0 || mode >= static_cast<int32_t>(ct.modes.size()) || !ct.modes[mode].active[b])
        throw std::invalid_argument("eda: tape: the element ran no such mode");
      const TapeMode& tm = ct.modes[mode];
      std::fill(ints, ints + 40, 0);
      std::fill(dbls, dbls + 34, 0.0);
      const ElemDev& el = tm.el[b];
      ints[0] = tm.w[b];
      ints[1] = tm.W[b];
      ints[2] = el.n_rungs;
      ints[3] = static_cast<int32_t>(el.used);
      ints[4] = static_cast<int32_t>(el.pts);
      for (int j = 0; j < el.n_rungs && j < 32; ++j) ints[5 + j] = el.rung_w[j];
      for (size_t j = 0; j < tm.ran.size(); ++j) {
        if (!tm.ran[j][b]) continue;
        ints[37] += 1;
        if (tm.acc[j][b] && j + 1 < 31) ints[38] |= 1 << (j + 1);
        if (j + 2 < 34) dbls[j + 2] = tm.thr[j][b];
      }
      ints[39] = ct.kexp[b];
      dbls[0] = tm.tol[b];
      dbls[1] = std::numeric_limits<double>::quiet_NaN();  // the amplitude bound is not kept
      const size_t item = run.dtype == EDA_F32 ? 4 : 8;
      const size_t n = static_cast<size_t>(ct.g.count);
      Context& ctx = context(run.device);
      EDA_HIP(hipStreamSynchronize(ctx.stream));
      if (residual != nullptr)
        EDA_HIP(hipMemcpy(residual, static_cast<const char*>(tm.r.p) + b * n * item, n * item, hipMemcpyDeviceToHost));
      if (L != nullptr)
        EDA_HIP(hipMemcpy(L, static_cast<const char*>(tm.L.p) + b * n * item, n * item, hipMemcpyDeviceToHost));
      if (lo != nullptr)
        EDA_HIP(hipMemcpy(lo, static_cast<const uint8_t*>(tm.lo.p) + b * n, n, hipMemcpyDeviceToHost));
      return;
    }
    throw std::invalid_argument("eda: tape: element out of range");
  });
}

int eda_timer_report(char* buf, size_t len, char* err, size_t errlen) {
  return guarded(err, errlen, [&] {
    const std::string js = timer_report();
    if (buf == nullptr || len <= js.size()) throw std::invalid_argument("eda: the report buffer is too small");
    std::memcpy(buf, js.c_str(), js.size() + 1);
  });
}

int eda_autotune_report(char* buf, size_t len, char* err, size_t errlen) {
  return guarded(err, errlen, [&] {
    const std::string js = autotune_report();
    if (buf == nullptr || len <= js.size()) throw std::invalid_argument("eda: the report buffer is too small");
    std::memcpy(buf, js.c_str(), js.size() + 1);
  });
}

}  // extern "C"

// ---- primitives ---------------------------------------------------------------------------------
namespace {
template <class F>
int prim(int32_t dtype, char* err, size_t errlen, F&& f) {
  return guarded(err, errlen, [&] {
    if (dtype != EDA_F32 && dtype != EDA_F64) throw std::invalid_argument("eda: dtype must be 0 or 1");
    int dev = 0;
    EDA_HIP(hipGetDevice(&dev));
    Context& ctx = context(dev);
    if (dtype == EDA_F32) f(float{}, ctx);
    else f(double{}, ctx);
    EDA_HIP(hipStreamSynchronize(ctx.stream));
    ctx.up.reset();
  });
}
}  // namespace

extern "C" {

int eda_prim_cube(int32_t dtype, int32_t dim, int64_t batch, const int64_t* n, int32_t op, const int32_t* win,
                  const void* in, void* out0, void* out1, char* err, size_t errlen) {
  return prim(dtype, err, errlen, [&](auto tv, Context& ctx) {
    using T = decltype(tv);
    const Geom g = PrimGeom<T>(dim, batch, n).g;
    if (op < 0 || op > 3) throw std::invalid_argument("eda: cube op must be 0..3");
    std::vector<int32_t> w(win, win + batch);
    int32_t mw = 0;
    for (int32_t v : w) {
      if (v != 0 && (v < 1 || v % 2 == 0)) throw std::invalid_argument("eda: windows must be odd (or 0)");
      mw = std::max(mw, v);
    }
    const int32_t* dw = ctx.up.put(w);
    enqueue_cube<T>(ctx.stream, g, static_cast<CubeOp>(op), static_cast<const T*>(in), static_cast<T*>(out0),
                    static_cast<T*>(out1), dw, mw);
  });
}

int eda_prim_stats(int32_t dtype, int32_t dim, int64_t batch, const int64_t* n, const void* f, double* energy,
                   double* amax, uint64_t* nonfinite, char* err, size_t errlen) {
  return prim(dtype, err, errlen, [&](auto tv, Context& ctx) {
    using T = decltype(tv);
    const Geom g = PrimGeom<T>(dim, batch, n).g;
    DevMem ws(elem_stats_ws_bytes(g) + 3 * 8 * batch + 1024, ctx.stream);
    double* de = static_cast<double*>(ws.p);
    uint64_t* da = reinterpret_cast<uint64_t*>(de + batch);
    uint64_t* dn = da + batch;
    char* rows = reinterpret_cast<char*>(dn + batch);
    enqueue_elem_stats<T>(ctx.stream, g, static_cast<const T*>(f), nullptr, de, da, dn, rows);
    std::vector<uint64_t> a(batch);
    EDA_HIP(hipMemcpyAsync(energy, de, 8 * batch, hipMemcpyDeviceToHost, ctx.stream));
    EDA_HIP(hipMemcpyAsync(a.data(), da, 8 * batch, hipMemcpyDeviceToHost, ctx.stream));
    EDA_HIP(hipMemcpyAsync(nonfinite, dn, 8 * batch, hipMemcpyDeviceToHost, ctx.stream));
    EDA_HIP(hipStreamSynchronize(ctx.stream));
    for (int64_t b = 0; b < batch; ++b) amax[b] = bits_d(a[b]);
  });
}

int eda_prim_select(int32_t dtype, int32_t dim, int64_t batch, const int64_t* n, const void* field,
                    const uint8_t* mask, int32_t bits, int32_t mode, double q, double* out, int64_t* count, char* err,
                    size_t errlen) {
  return prim(dtype, err, errlen, [&](auto tv, Context& ctx) {
    using T = decltype(tv);
    const Geom g = PrimGeom<T>(dim, batch, n).g;
    if (mode != 0 && !(q >= 0.0 && q <= 1.0)) throw std::invalid_argument("eda: quantile: q must lie in [0, 1]");
    DevMem ws(select_ws_bytes(g) + 16 * batch + 1024, ctx.stream);
    double* dout = reinterpret_cast<double*>(static_cast<char*>(ws.p) + select_ws_bytes(g));
    int64_t* dcnt = reinterpret_cast<int64_t*>(dout + batch);
    Selection<T> sel{};
    sel.src = SelSrc::kField;
    sel.field = static_cast<const T*>(field);
    sel.mask = mask;
    sel.bits = static_cast<uint8_t>(bits);
    enqueue_select<T>(ctx.stream, g, sel, mode, q, dout, dcnt, nullptr, ws.p);
    EDA_HIP(hipMemcpyAsync(out, dout, 8 * batch, hipMemcpyDeviceToHost, ctx.stream));
    if (count) EDA_HIP(hipMemcpyAsync(count, dcnt, 8 * batch, hipMemcpyDeviceToHost, ctx.stream));
    EDA_HIP(hipStreamSynchronize(ctx.stream));
  });
}

int eda_prim_noise(int32_t dtype, int32_t dim, int64_t batch, const int64_t* n, const void* r, int32_t per_axis,
                   double* out, char* err, size_t errlen) {
  return prim(dtype, err, errlen, [&](auto tv, Context& ctx) {
    using T = decltype(tv);
    const Geom g = PrimGeom<T>(dim, batch, n).g;
    DevMem ws(noise_ws_bytes(g) + 8 * batch + 1024, ctx.stream);
    double* dout = reinterpret_cast<double*>(static_cast<char*>(ws.p) + noise_ws_bytes(g));
    enqueue_noise_sigma<T>(ctx.stream, g, static_cast<const T*>(r), per_axis != 0, dout, nullptr, ws.p);
    EDA_HIP(hipMemcpyAsync(out, dout, 8 * batch, hipMemcpyDeviceToHost, ctx.stream));
    EDA_HIP(hipStreamSynchronize(ctx.stream));
  });
}

int eda_prim_tau(int32_t dtype, int32_t dim, int64_t batch, const int64_t* n, const void* r, const int32_t* Wd,
                 const int32_t* Wl, const double* sigma_n, const double* tol, void* tau, int32_t* nonfinite, char* err,
                 size_t errlen) {
  return prim(dtype, err, errlen, [&](auto tv, Context& ctx) {
    using T = decltype(tv);
    const Geom g = PrimGeom<T>(dim, batch, n).g;
    const MethodCtl c = method_ctl(g, std::is_same_v<T, float>);
    DevMem tmp(static_cast<size_t>(g.batch * g.count) * sizeof(T) + 4 * batch + 1024, ctx.stream);
    T* dtmp = static_cast<T*>(tmp.p);
    int32_t* dnf = reinterpret_cast<int32_t*>(dtmp + g.batch * g.count);
    std::vector<int32_t> wd(Wd, Wd + batch), wl(Wl, Wl + batch);
    const int32_t mWd = *std::max_element(wd.begin(), wd.end()), mWl = *std::max_element(wl.begin(), wl.end());
    const int32_t* dWd = ctx.up.put(wd);
    const int32_t* dWl = ctx.up.put(wl);
    const double* dsig = ctx.up.put(sigma_n, batch);
    const double* dtol = ctx.up.put(tol, batch);
    EDA_HIP(hipMemsetAsync(dnf, 0, 4 * batch, ctx.stream));
    const T* rp = static_cast<const T*>(r);
    T* tp = static_cast<T*>(tau);
    enqueue_cube<T>(ctx.stream, g, CubeOp::kBox, rp, dtmp, nullptr, dWd, mWd);
    each(ctx.stream, g, nullptr, [=] __device__(int64_t b, int64_t i) {
      const int64_t at = b * g.count + i;
      const T det = rp[at] - dtmp[at];
      tp[at] = det * det;
    });
    enqueue_cube<T>(ctx.stream, g, CubeOp::kBox, tp, tp, nullptr, dWl, mWl);
    const double theta = c.sig_theta, k_noise = c.sig_noise;
    each(ctx.stream, g, nullptr, [=] __device__(int64_t b, int64_t i) {
      const int64_t at = b * g.count + i;
      const T be = tp[at];
      if (!(isfinite(be) && isfinite(dsig[b]))) dnf[b] = 1;
      const T lev = sqrt(be > T(0) ? be : T(0));
      const double f = k_noise * dsig[b];
      const double fl = f > 0.0 ? f : 0.0;
      const double ld = static_cast<double>(lev);
      double t = theta * (ld > fl ? ld : fl);
      t = t > dtol[b] ? t : dtol[b];
      tp[at] = static_cast<T>(t);
    });
    EDA_HIP(hipMemcpyAsync(nonfinite, dnf, 4 * batch, hipMemcpyDeviceToHost, ctx.stream));
    EDA_HIP(hipStreamSynchronize(ctx.stream));
  });
}

int eda_prim_flags(int32_t dtype, int32_t dim, int64_t batch, const int64_t* n, const void* r, const void* tau,
                   uint8_t* flags, uint64_t* ext, char* err, size_t errlen) {
  return prim(dtype, err, errlen, [&](auto tv, Context& ctx) {
    using T = decltype(tv);
    const Geom g = PrimGeom<T>(dim, batch, n).g;
    DevMem tmp(static_cast<size_t>(g.batch * g.count) * sizeof(T) + 24 * batch + 1024, ctx.stream);
    T* scratch = static_cast<T*>(tmp.p);
    uint64_t* dext = reinterpret_cast<uint64_t*>(scratch + g.batch * g.count);
    EDA_HIP(hipMemsetAsync(dext, 0, 24 * batch, ctx.stream));
    enqueue_sig_flags<T>(ctx.stream, g, static_cast<const T*>(r), static_cast<const T*>(tau), flags, dext, nullptr,
                         scratch);
    EDA_HIP(hipMemcpyAsync(ext, dext, 24 * batch, hipMemcpyDeviceToHost, ctx.stream));
    EDA_HIP(hipStreamSynchronize(ctx.stream));
  });
}

int eda_prim_raw_window(int32_t dtype, int32_t dim, int64_t batch, const int64_t* n, const uint8_t* flags,
                        const int32_t* win, void* om, char* err, size_t errlen) {
  return prim(dtype, err, errlen, [&](auto tv, Context& ctx) {
    using T = decltype(tv);
    const Geom g = PrimGeom<T>(dim, batch, n).g;
    const MethodCtl c = method_ctl(g, std::is_same_v<T, float>);
    DevMem tmp(static_cast<size_t>(g.batch * g.count) * sizeof(T), ctx.stream);
    T* dtmp = static_cast<T*>(tmp.p);
    std::vector<int32_t> w(win, win + batch);
    const int32_t mw = *std::max_element(w.begin(), w.end());
    const int32_t* dw = ctx.up.put(w);
    T* o = static_cast<T*>(om);
    const T two_kappa = static_cast<T>(2.0 * c.kappa);
    const T tiny = static_cast<T>(c.tiny);
    for (int a = 0; a < g.dim; ++a) {
      each(ctx.stream, g, nullptr, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        dtmp[at] = ((flags[at] >> a) & 1u) ? T(1) : T(0);
      });
      enqueue_cube<T>(ctx.stream, g, CubeOp::kBox, dtmp, dtmp, nullptr, dw, mw);
      const bool last = a == g.dim - 1;
      each(ctx.stream, g, nullptr, [=] __device__(int64_t b, int64_t i) {
        const int64_t at = b * g.count + i;
        const T sq = dtmp[at] * dtmp[at];
        T acc = a == 0 ? sq : o[at] + sq;
        if (last) {
          const T dens = sqrt(acc);
          acc = dens > T(0) ? two_kappa / (dens > tiny ? dens : tiny) : inf_t<T>();
        }
        o[at] = acc;
      });
    }
  });
}

int eda_prim_ccl(int32_t dim, int64_t batch, const int64_t* n, const uint8_t* mask, int32_t bits, int32_t* labels,
                 char* err, size_t errlen) {
  return prim(EDA_F32, err, errlen, [&](auto, Context& ctx) {
    const Geom g = PrimGeom<float>(dim, batch, n).g;
    enqueue_ccl(ctx.stream, g, mask, static_cast<uint8_t>(bits), labels, nullptr);
  });
}

int eda_prim_dct(int32_t dtype, int32_t dim, int64_t batch, const int64_t* n, const void* in, void* out,
                 int32_t inverse, int32_t odd, char* err, size_t errlen) {
  return prim(dtype, err, errlen, [&](auto tv, Context& ctx) {
    using T = decltype(tv);
    const Geom g = PrimGeom<T>(dim, batch, n).g;
    DevMem tmp(static_cast<size_t>(g.batch * g.count) * sizeof(T), ctx.stream);
    if (in == out) throw std::invalid_argument("eda: dct: in and out must differ");
    enqueue_dct<T>(ctx.stream, ctx.device, g, static_cast<const T*>(in), static_cast<T*>(out), static_cast<T*>(tmp.p),
                   inverse, odd);
  });
}

int eda_prim_orientation(int32_t dtype, int32_t dim, int64_t batch, const int64_t* n, const void* r, const int32_t* w,
                         void* nvec, void* coh, char* err, size_t errlen) {
  return prim(dtype, err, errlen, [&](auto tv, Context& ctx) {
    using T = decltype(tv);
    if (dim < 2) throw std::invalid_argument("eda: orientation needs 2 or 3 axes");
    const Geom g = PrimGeom<T>(dim, batch, n).g;
    const int K = dim * (dim + 1) / 2;
    const size_t fbytes = static_cast<size_t>(g.batch * g.count) * sizeof(T) * (3 + K);
    const size_t dbytes = static_cast<size_t>(orient_scratch_doubles(batch)) * sizeof(double);
    DevMem work(fbytes + dbytes + 4 * batch + 1024, ctx.stream);
    double* dscr = reinterpret_cast<double*>(static_cast<char*>(work.p) + (fbytes + 255) / 256 * 256);
    int32_t* dnf = reinterpret_cast<int32_t*>(dscr + orient_scratch_doubles(batch));
    std::vector<int32_t> wv(w, w + batch);
    const int32_t* dw = ctx.up.put(wv);
    enqueue_orientation<T>(ctx.stream, ctx.device, g, static_cast<const T*>(r), dw, nullptr, static_cast<T*>(nvec),
                           static_cast<T*>(coh), static_cast<T*>(work.p), dnf, dscr);
  });
}

}  // extern "C"
