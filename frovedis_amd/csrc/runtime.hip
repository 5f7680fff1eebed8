// Runtime pieces of libfrovedis_hip: last-error text, the device workspace
// pool, pinned host staging and the init/finalize entry points.
#include "fhip_common.hpp"

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

int fhip_dev_free(void* p) {
  return fhip::guarded([&] {
    if (p) FHIP_CHECK(hipFree(p));
  });
}

}  // extern "C"
