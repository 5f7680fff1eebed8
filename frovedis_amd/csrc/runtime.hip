// Runtime pieces of libfrovedis_hip: last-error text, the device workspace
// pool, pinned host staging and the init/finalize entry points.
#include "fhip_common.hpp"

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>

namespace fhip {

namespace {
thread_local std::string g_last_error;
constexpr size_t kPoolAlign = size_t(2) << 20;  // 2 MiB granules
}  // namespace

void set_last_error(const std::string& msg) { g_last_error = msg; }

Pool& Pool::get() {
  static Pool* p = new Pool();  // intentionally leaked: outlives static dtors
  return *p;
}

void* Pool::alloc(size_t bytes) {
  size_t sz = ((bytes + kPoolAlign - 1) / kPoolAlign) * kPoolAlign;
  int dev = 0;
  FHIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu_);
  // best fit among this device's cached blocks up to 2x the request
  auto it = free_.lower_bound(Key{dev, sz});
  if (it != free_.end() && it->first.dev == dev && it->first.size <= 2 * sz) {
    void* p = it->second;
    const Key have = it->first;
    free_.erase(it);
    live_[p] = have;
    return p;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, sz);
  if (e != hipSuccess) {
    // drop this device's cache and retry once
    (void)hipGetLastError();
    FHIP_CHECK(hipDeviceSynchronize());
    for (auto i = free_.begin(); i != free_.end();) {
      if (i->first.dev != dev) { ++i; continue; }
      (void)hipFree(i->second);
      cached_ -= i->first.size;
      i = free_.erase(i);
    }
    e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error(FHIP_ERR_NOMEM, "device allocation of " + std::to_string(sz) +
                                      " bytes failed");
    }
  }
  live_[p] = Key{dev, sz};
  cached_ += sz;
  return p;
}

void Pool::release(void* p) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = live_.find(p);
  if (it == live_.end()) return;
  free_.emplace(it->second, p);
  live_.erase(it);
}

void Pool::trim() {
  int cur = 0;
  FHIP_CHECK(hipGetDevice(&cur));
  std::lock_guard<std::mutex> lk(mu_);
  for (auto& kv : free_) {
    // each block is freed with its own device current
    if (kv.first.dev != cur) (void)hipSetDevice(kv.first.dev);
    (void)hipFree(kv.second);
    if (kv.first.dev != cur) (void)hipSetDevice(cur);
    cached_ -= kv.first.size;
  }
  free_.clear();
}

namespace {
struct StreamState {
  bool used = false;
  hipStream_t last = nullptr;
  hipEvent_t ev = nullptr;
};
std::mutex g_stream_mu;
std::map<int, StreamState>& stream_states() {
  static auto* m = new std::map<int, StreamState>();  // leaked: outlives static dtors
  return *m;
}
}  // namespace

void stream_order(hipStream_t st) {
  int dev = 0;
  FHIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_stream_mu);
  StreamState& s = stream_states()[dev];
  if (s.used && s.last != st) {
    if (!s.ev) FHIP_CHECK(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
    FHIP_CHECK(hipEventRecord(s.ev, s.last));
    FHIP_CHECK(hipStreamWaitEvent(st, s.ev, 0));
  }
  s.used = true;
  s.last = st;
}

HostStage::HostStage(size_t bytes) {
  if (bytes == 0) return;
  FHIP_CHECK(hipHostMalloc(&p_, bytes, hipHostMallocDefault));
}
HostStage::~HostStage() {
  if (p_) (void)hipHostFree(p_);
}

namespace {
constexpr int kPinSlots = 4;
struct PinSlots {
  void* p[kPinSlots] = {};
  size_t cap[kPinSlots] = {};
  hipEvent_t ev = nullptr;
};
// per thread, intentionally leaked (with its pinned buffers and event): it
// must outlive static destructors that run after the HIP runtime has gone
thread_local PinSlots* g_pin = nullptr;
}  // namespace

void* pinned_scratch(int slot, size_t bytes) {
  FHIP_REQUIRE(slot >= 0 && slot < kPinSlots, FHIP_ERR_ARG, "pinned_scratch: bad slot");
  if (!g_pin) g_pin = new PinSlots();
  if (g_pin->cap[slot] < bytes) {
    size_t c = size_t(64) << 10;
    while (c < bytes) c *= 2;
    if (g_pin->p[slot]) {
      (void)hipHostFree(g_pin->p[slot]);
      g_pin->p[slot] = nullptr;
      g_pin->cap[slot] = 0;
    }
    FHIP_CHECK(hipHostMalloc(&g_pin->p[slot], c, hipHostMallocDefault));
    g_pin->cap[slot] = c;
  }
  return g_pin->p[slot];
}

namespace {
// two pinned chunks and their events per thread (leaked like the slots)
constexpr size_t kStageChunk = size_t(16) << 20;
constexpr size_t kStageDirect = size_t(1) << 20;  // smaller copies go straight
struct Staging {
  void* buf[2] = {};
  hipEvent_t ev[2] = {};
};
thread_local Staging* g_stage = nullptr;

Staging& staging() {
  if (!g_stage) {
    g_stage = new Staging();
    for (int i = 0; i < 2; i++) {
      FHIP_CHECK(hipHostMalloc(&g_stage->buf[i], kStageChunk, hipHostMallocDefault));
      FHIP_CHECK(hipEventCreateWithFlags(&g_stage->ev[i], hipEventDisableTiming));
      FHIP_CHECK(hipEventRecord(g_stage->ev[i], nullptr));
    }
  }
  return *g_stage;
}
}  // namespace

void copy_h2d(void* dst, const void* src, size_t bytes, hipStream_t st) {
  if (bytes == 0) return;
  if (bytes <= kStageDirect) {
    FHIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    FHIP_CHECK(hipStreamSynchronize(st));
    return;
  }
  Staging& s = staging();
  char* d = static_cast<char*>(dst);
  const char* h = static_cast<const char*>(src);
  int b = 0;
  for (size_t off = 0; off < bytes; off += kStageChunk, b ^= 1) {
    const size_t len = std::min(kStageChunk, bytes - off);
    FHIP_CHECK(hipEventSynchronize(s.ev[b]));  // the chunk's previous DMA is done
    std::memcpy(s.buf[b], h + off, len);
    FHIP_CHECK(hipMemcpyAsync(d + off, s.buf[b], len, hipMemcpyHostToDevice, st));
    FHIP_CHECK(hipEventRecord(s.ev[b], st));
  }
  FHIP_CHECK(hipStreamSynchronize(st));
}

void copy_d2h(void* dst, const void* src, size_t bytes, hipStream_t st) {
  if (bytes == 0) return;
  if (bytes <= kStageDirect) {
    FHIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    FHIP_CHECK(hipStreamSynchronize(st));
    return;
  }
  Staging& s = staging();
  char* h = static_cast<char*>(dst);
  const char* d = static_cast<const char*>(src);
  const size_t chunks = (bytes + kStageChunk - 1) / kStageChunk;
  auto issue = [&](size_t c) {
    const size_t off = c * kStageChunk, len = std::min(kStageChunk, bytes - off);
    FHIP_CHECK(hipMemcpyAsync(s.buf[c & 1], d + off, len, hipMemcpyDeviceToHost, st));
    FHIP_CHECK(hipEventRecord(s.ev[c & 1], st));
  };
  FHIP_CHECK(hipStreamSynchronize(st));  // earlier users of both chunks are done
  issue(0);
  for (size_t c = 0; c < chunks; c++) {
    if (c + 1 < chunks) issue(c + 1);  // its buffer was emptied in step c - 1
    FHIP_CHECK(hipEventSynchronize(s.ev[c & 1]));
    const size_t off = c * kStageChunk, len = std::min(kStageChunk, bytes - off);
    std::memcpy(h + off, s.buf[c & 1], len);
  }
}

hipEvent_t thread_event() {
  if (!g_pin) g_pin = new PinSlots();
  if (!g_pin->ev) FHIP_CHECK(hipEventCreateWithFlags(&g_pin->ev, hipEventDisableTiming));
  return g_pin->ev;
}

namespace {
struct ProfEntry {
  std::string name;
  hipEvent_t a, b;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<ProfEntry> g_prof;
// recorded events go back to this list when the records are cleared, so a
// timed region creates no events (enabling fills it with kProfEvents)
std::vector<hipEvent_t> g_prof_free;
constexpr size_t kProfEvents = 512;
std::vector<std::string> g_prof_only;  // names recorded (empty: all)
}  // namespace

bool prof_wanted(const char* name) {
  if (!g_prof_on) return false;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (g_prof_only.empty()) return true;
  for (auto& s : g_prof_only)
    if (s == name) return true;
  return false;
}

hipEvent_t prof_event() {
  {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (!g_prof_free.empty()) {
      hipEvent_t e = g_prof_free.back();
      g_prof_free.pop_back();
      return e;
    }
  }
  hipEvent_t e = nullptr;
  FHIP_CHECK(hipEventCreate(&e));
  return e;
}

bool prof_enabled() { return g_prof_on; }

// ---- tunables ---------------------------------------------------------------
namespace {
std::mutex g_tune_mu;
std::map<std::string, long long>& tune_map() {
  static auto* m = new std::map<std::string, long long>();
  return *m;
}
}  // namespace

long long tune_get(const char* name, long long dflt) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = tune_map().find(name);
  return it == tune_map().end() ? dflt : it->second;
}

void prof_record(const char* name, hipEvent_t a, hipEvent_t b) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof.push_back(ProfEntry{name, a, b});
}

static void prof_clear_locked() {
  for (auto& e : g_prof) {
    (void)hipEventSynchronize(e.b);  // both may still be pending on their stream
    g_prof_free.push_back(e.a);
    g_prof_free.push_back(e.b);
  }
  g_prof.clear();
}

int fill_blocks(int threads_per_block, int blocks_per_cu) {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    FHIP_CHECK(hipGetDevice(&dev));
    FHIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (cus <= 0) cus = 256;
  }
  (void)threads_per_block;
  return cus * blocks_per_cu;
}

}  // namespace fhip

extern "C" {

const char* fhip_last_error(void) { return fhip::g_last_error.c_str(); }

int fhip_init(int device) {
  return fhip::guarded([&] {
    FHIP_CHECK(hipSetDevice(device));
    FHIP_CHECK(hipFree(nullptr));
  });
}

int fhip_finalize(void) {
  return fhip::guarded([&] {
    FHIP_CHECK(hipDeviceSynchronize());
    fhip::Pool::get().trim();
  });
}

int fhip_prof_enable(int on) {
  return fhip::guarded([&] {
    std::lock_guard<std::mutex> lk(fhip::g_prof_mu);
    fhip::g_prof_on = on != 0;
    fhip::prof_clear_locked();
    while (on && fhip::g_prof_free.size() < fhip::kProfEvents) {
      hipEvent_t e = nullptr;
      FHIP_CHECK(hipEventCreate(&e));
      fhip::g_prof_free.push_back(e);
    }
  });
}

int fhip_prof_filter(const char* names) {
  return fhip::guarded([&] {
    std::lock_guard<std::mutex> lk(fhip::g_prof_mu);
    fhip::g_prof_only.clear();
    if (!names) return;
    std::string all(names);
    size_t i = 0;
    while (i <= all.size()) {
      size_t j = all.find(',', i);
      if (j == std::string::npos) j = all.size();
      if (j > i) fhip::g_prof_only.push_back(all.substr(i, j - i));
      i = j + 1;
    }
  });
}

int fhip_prof_query(const char* name, double* total_ms, size_t* launches) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(name && total_ms && launches, FHIP_ERR_ARG, "prof_query: null pointer");
    std::lock_guard<std::mutex> lk(fhip::g_prof_mu);
    double t = 0;
    size_t c = 0;
    for (auto& e : fhip::g_prof) {
      if (e.name != name) continue;
      FHIP_CHECK(hipEventSynchronize(e.b));
      float ms = 0;
      FHIP_CHECK(hipEventElapsedTime(&ms, e.a, e.b));
      t += ms;
      c++;
    }
    *total_ms = t;
    *launches = c;
  });
}

int fhip_tune_set(const char* name, long long value) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(name, FHIP_ERR_ARG, "tune_set: null name");
    std::lock_guard<std::mutex> lk(fhip::g_tune_mu);
    if (value < 0) fhip::tune_map().erase(name);
    else fhip::tune_map()[name] = value;
  });
}

int fhip_tune_get(const char* name, long long dflt, long long* value) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(name && value, FHIP_ERR_ARG, "tune_get: null pointer");
    *value = fhip::tune_get(name, dflt);
  });
}

// The tunables as one flat JSON object {"name": integer, ...}.
int fhip_autotune_save(const char* path) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(path, FHIP_ERR_ARG, "autotune_save: null path");
    std::lock_guard<std::mutex> lk(fhip::g_tune_mu);
    FILE* f = std::fopen(path, "w");
    FHIP_REQUIRE(f, FHIP_ERR_ARG, std::string("autotune_save: cannot open ") + path);
    std::fputs("{", f);
    bool first = true;
    for (auto& kv : fhip::tune_map()) {
      std::fprintf(f, "%s\n  \"%s\": %lld", first ? "" : ",", kv.first.c_str(), kv.second);
      first = false;
    }
    std::fputs("\n}\n", f);
    std::fclose(f);
  });
}

int fhip_autotune_load(const char* path) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(path, FHIP_ERR_ARG, "autotune_load: null path");
    FILE* f = std::fopen(path, "r");
    FHIP_REQUIRE(f, FHIP_ERR_ARG, std::string("autotune_load: cannot open ") + path);
    std::string text;
    char buf[4096];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, got);
    std::fclose(f);
    // "name": integer pairs; anything else is a format error
    std::map<std::string, long long> parsed;
    size_t i = text.find('{');
    FHIP_REQUIRE(i != std::string::npos, FHIP_ERR_ARG, "autotune_load: not a JSON object");
    for (;;) {
      size_t q0 = text.find('"', i + 1);
      size_t close = text.find('}', i + 1);
      if (q0 == std::string::npos || (close != std::string::npos && close < q0)) break;
      size_t q1 = text.find('"', q0 + 1);
      size_t colon = text.find(':', q1 + 1);
      FHIP_REQUIRE(q1 != std::string::npos && colon != std::string::npos, FHIP_ERR_ARG,
                   "autotune_load: malformed entry");
      char* end = nullptr;
      const long long v = std::strtoll(text.c_str() + colon + 1, &end, 10);
      FHIP_REQUIRE(end != text.c_str() + colon + 1, FHIP_ERR_ARG,
                   "autotune_load: value of " + text.substr(q0 + 1, q1 - q0 - 1) +
                       " is not an integer");
      parsed[text.substr(q0 + 1, q1 - q0 - 1)] = v;
      i = size_t(end - text.c_str());
    }
    std::lock_guard<std::mutex> lk(fhip::g_tune_mu);
    for (auto& kv : parsed) fhip::tune_map()[kv.first] = kv.second;
  });
}

int fhip_dev_free(void* p) {
  return fhip::guarded([&] {
    if (p) FHIP_CHECK(hipFree(p));
  });
}

}  // extern "C"
