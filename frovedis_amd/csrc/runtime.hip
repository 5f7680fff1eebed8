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
  std::lock_guard<std::mutex> lk(mu_);
  // best fit among cached blocks up to 2x the request
  auto it = free_.lower_bound(sz);
  if (it != free_.end() && it->first <= 2 * sz) {
    void* p = it->second;
    size_t have = it->first;
    free_.erase(it);
    live_[p] = have;
    return p;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, sz);
  if (e != hipSuccess) {
    // drop the cache and retry once
    (void)hipGetLastError();
    FHIP_CHECK(hipDeviceSynchronize());
    for (auto& kv : free_) { (void)hipFree(kv.second); cached_ -= kv.first; }
    free_.clear();
    e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error(FHIP_ERR_NOMEM, "device allocation of " + std::to_string(sz) +
                                      " bytes failed");
    }
  }
  live_[p] = sz;
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
  std::lock_guard<std::mutex> lk(mu_);
  for (auto& kv : free_) { (void)hipFree(kv.second); cached_ -= kv.first; }
  free_.clear();
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
