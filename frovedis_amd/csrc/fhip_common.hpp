// Shared internals of libfrovedis_hip: status codes, error text, type
// codes, the device workspace pool and small device helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/frovedis_hip.h"

namespace fhip {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define FHIP_CHECK(expr)                                                      \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess)                                                     \
      throw ::fhip::Error(FHIP_ERR_HIP, std::string(#expr) + ": " +           \
                                            hipGetErrorString(e_));           \
  } while (0)

#define FHIP_REQUIRE(cond, code, msg)                                         \
  do {                                                                        \
    if (!(cond)) throw ::fhip::Error((code), (msg));                          \
  } while (0)

// Runs `body` and converts exceptions into status codes (the C ABI never
// lets a C++ exception escape).
template <class F>
int guarded(F&& body) {
  try {
    body();
    return FHIP_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return FHIP_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return FHIP_ERR_INTERNAL;
  }
}

// ---------------------------------------------------------------------------
// type codes (same numbering as the C header)
// ---------------------------------------------------------------------------
inline int type_bytes(int t) {
  switch (t) {
    case FHIP_I32: case FHIP_U32: case FHIP_F32: return 4;
    case FHIP_I64: case FHIP_U64: case FHIP_F64: return 8;
    default: throw Error(FHIP_ERR_TYPE, "unsupported type code " + std::to_string(t));
  }
}
inline bool type_is_signed_int(int t) { return t == FHIP_I32 || t == FHIP_I64; }
inline bool type_is_float(int t) { return t == FHIP_F32 || t == FHIP_F64; }

// ---------------------------------------------------------------------------
// device workspace: a grow-only cache of device blocks keyed by (device,
// size class).  Blocks are handed out for the duration of one library call
// and go back to the cache when the call returns, while its kernels may
// still be queued.  Reuse is safe because every entry point orders its
// stream after the previous call's stream on that device (stream_order
// below), so a block is only handed to work that runs after all the work of
// the call that released it.  Calls on one device are serialised by the
// caller (one host thread per rank, as the reference's
// MPI_THREAD_SERIALIZED ranks are).
// ---------------------------------------------------------------------------
class Pool {
 public:
  static Pool& get();
  void* alloc(size_t bytes);
  void release(void* p);
  void trim();  // frees every cached block (device must be idle)
  size_t cached_bytes() const { return cached_; }

 private:
  struct Key {
    int dev;
    size_t size;
    bool operator<(const Key& o) const { return dev != o.dev ? dev < o.dev : size < o.size; }
  };
  std::mutex mu_;
  std::multimap<Key, void*> free_;
  std::map<void*, Key> live_;
  size_t cached_ = 0;
};

// Makes stream `st` wait for everything queued so far on the stream of the
// previous library call on the current device (when that was another
// stream), then records `st` as the device's last stream.
void stream_order(hipStream_t st);

// RAII handle over a pool block.
class DBuf {
 public:
  DBuf() = default;
  explicit DBuf(size_t bytes) { reset(bytes); }
  ~DBuf() { free(); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) { free(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
    return *this;
  }
  void reset(size_t bytes) {
    free();
    if (bytes) { p_ = Pool::get().alloc(bytes); n_ = bytes; }
  }
  void free() {
    if (p_) Pool::get().release(p_);
    p_ = nullptr; n_ = 0;
  }
  void* get() const { return p_; }
  template <class T> T* as() const { return static_cast<T*>(p_); }
  size_t size() const { return n_; }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

// Pinned host staging (hipHostMalloc), for the host-pointer entry points.
class HostStage {
 public:
  explicit HostStage(size_t bytes);
  ~HostStage();
  void* get() const { return p_; }
 private:
  void* p_ = nullptr;
};

// Pinned host memory for small stream-ordered copies: buffer `slot` (0 to
// 3) of the calling thread, at least `bytes` long.  A thread's buffers grow
// on demand and are kept until the process ends; a caller reuses a slot only
// once the copies that used it have completed.
void* pinned_scratch(int slot, size_t bytes);
// An event of the calling thread (no timing), for host waits on a point of
// a stream; kept until the process ends.
hipEvent_t thread_event();

// The stream of an entry point, ordered after the previous call's stream.
inline hipStream_t as_stream(void* s) {
  hipStream_t st = static_cast<hipStream_t>(s);
  stream_order(st);
  return st;
}

// Host <-> device copies of host-pointer entry points, staged through two
// pinned buffers of the calling thread so that the host copy of one chunk
// overlaps the DMA of the other.  Both return once the data has arrived.
void copy_h2d(void* dst, const void* src, size_t bytes, hipStream_t st);
void copy_d2h(void* dst, const void* src, size_t bytes, hipStream_t st);

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
constexpr int kWave = 64;

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

// inclusive wave prefix sum (64 lanes)
template <class T>
__device__ __forceinline__ T wave_incl_scan(T x) {
  const unsigned lane = __lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T y = __shfl_up(x, o, 64);
    if (lane >= unsigned(o)) x += y;
  }
  return x;
}

template <class T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// Exclusive prefix over a block of NT threads; `tmp` holds NT/64 entries in
// LDS.  Returns the exclusive prefix; *total gets the block total.
template <int NT, class T>
__device__ __forceinline__ T block_excl_scan(T v, T* tmp, T* total) {
  const unsigned lane = __lane_id();
  const unsigned wave = threadIdx.x / 64;
  T x = wave_incl_scan(v);
  if (lane == 63) tmp[wave] = x;
  __syncthreads();
  T before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; w++) {
    T t = tmp[w];
    if (unsigned(w) < wave) before += t;
    all += t;
  }
  __syncthreads();
  if (total) *total = all;
  return before + x - v;
}

inline size_t ceil_div(size_t a, size_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// kernel timing (off by default): HIP events recorded on the launch stream
// around a named kernel; read back through fhip_prof_query.
// ---------------------------------------------------------------------------
bool prof_enabled();
bool prof_wanted(const char* name);  // enabled, and the name passes the filter
hipEvent_t prof_event();  // from the recycled events when there are any
void prof_record(const char* name, hipEvent_t start, hipEvent_t stop);

class ProfScope {
 public:
  ProfScope(const char* name, hipStream_t st) : name_(name), st_(st) {
    if (!prof_wanted(name)) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st_, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    a_ = prof_event();
    b_ = prof_event();
    FHIP_CHECK(hipEventRecord(a_, st_));
  }
  ~ProfScope() {
    if (!a_) return;
    if (hipEventRecord(b_, st_) == hipSuccess) prof_record(name_, a_, b_);
  }
  ProfScope(const ProfScope&) = delete;
  ProfScope& operator=(const ProfScope&) = delete;

 private:
  const char* name_;
  hipStream_t st_;
  hipEvent_t a_ = nullptr, b_ = nullptr;
};

// Tunable parameters (set by the autotuner through fhip_tune_set); the
// default is returned when a name was never set.
long long tune_get(const char* name, long long dflt);

// number of workgroups that fill the device for grid-stride kernels
int fill_blocks(int threads_per_block, int blocks_per_cu = 8);

// The heavy keys of a column known in advance (from a model of the column):
// key bits (sign-extended for signed keys) and their counts, the heaviest
// first, counted over `total` rows.
struct HeavyKeys;

// The heavy-key table placed from a HeavyKeys list, kept by the list's owner
// (a column model) so that repeated group-bys do not place it again.  The
// table depends on the list and on the slot count and coverage rule in force
// (both tunables), which are the cache key.
struct HotPlanCache {
  std::mutex mu;
  bool valid = false;
  int Sh = 0;               // slots (0: no plan)
  long long min_pct = -1;   // coverage rule the plan was made with
  bool used = false;        // the plan passed the coverage rule
  std::vector<unsigned long long> table;
};

struct HeavyKeys {
  const unsigned long long* keys = nullptr;
  const unsigned long long* counts = nullptr;
  size_t n = 0;
  size_t total = 0;
  HotPlanCache* cache = nullptr;  // optional, owned by the caller
};

// The fused hash group-by + aggregation (hashagg.hip); heavy (optional)
// replaces the sample the heavy-key table is planned from.
void group_by_agg(int ktype, const void* keys, size_t n, const fhip_agg* aggs, int naggs,
                  void* keys_out, size_t capacity, size_t* ngroups, size_t table_hint,
                  hipStream_t st, const HeavyKeys* heavy = nullptr, bool narrow_cold = true);

}  // namespace fhip
