// unique_hashtable<K, size_t> for gfx950 (dataframe/hashtable.hpp).
//
// The table keeps the reference's layout so that every observable result
// (lookups, miss lists, check_existence, is_unique_ok and the all_keys order)
// is the same:
//   - table_size: smallest listed prime >= 3n (:23, :115-152, :1174-1179);
//   - home slot: unsigned key bits % table_size (:1190-1212);
//   - the last row in index order to hash to a slot owns it (atomicMax on
//     row + 1 replaces the reference's sequential overwrite, :1197-1200);
//   - rows whose key differs from their slot owner's key form the conflict
//     list in row order, stable radix-sorted by key (:1216-1235);
//   - probe: slot key, else the first equal entry of the conflict list
//     (lower_bound, :1442-1515).
// In HBM: one 32-byte slot per table entry holding owner, key and value
// (a probe reads one sector), the sorted conflict arrays, and for integer
// keys a secondary open-addressing index over the distinct conflict keys
// that answers "first equal entry" in one or two sector reads instead of a
// binary search of ~log2(conflicts) dependent reads.
#include "hashtable.hpp"

#include "fhip_device.hpp"
#include "radix_sort.hpp"

namespace fhip {
namespace {

const unsigned long long kPrimes[] = {
    8 + 3, 16 + 3, 32 + 5, 64 + 3, 128 + 3, 256 + 27, 512 + 9, 1024 + 9, 2048 + 5, 4096 + 3,
    8192 + 27, 16384 + 43, 32768 + 3, 65536 + 45, 131072 + 29, 262144 + 3, 524288 + 21,
    1048576 + 7, 2097152 + 17, 4194304 + 15, 8388608 + 9, 16777216 + 43, 33554432 + 35,
    67108864 + 15, 134217728 + 29, 268435456 + 3, 536870912 + 11, 1073741824 + 85,
    2147483648ull + 11, 4294967296ull + 15, 8589934592ull + 17, 17179869184ull + 25,
    34359738368ull + 53};

template <class K>
__device__ __forceinline__ unsigned long long home_slot(K k, unsigned long long ts) {
  if constexpr (sizeof(K) == 4) {
    // uint32 bits promoted to size_t, as the reference does (:1190-1197)
    unsigned b;
    __builtin_memcpy(&b, &k, 4);
    return ts <= 0xFFFFFFFFull ? (unsigned long long)(b % unsigned(ts)) : (unsigned long long)b;
  } else {
    unsigned long long b;
    __builtin_memcpy(&b, &k, 8);
    return b % ts;
  }
}

template <class K>
__device__ __forceinline__ unsigned long long key_to_bits(K k) {
  if constexpr (sizeof(K) == 4) {
    unsigned b;
    __builtin_memcpy(&b, &k, 4);
    return b;
  } else {
    unsigned long long b;
    __builtin_memcpy(&b, &k, 8);
    return b;
  }
}

template <class K>
__device__ __forceinline__ K bits_to_key(unsigned long long b) {
  K k;
  if constexpr (sizeof(K) == 4) {
    const unsigned lo = unsigned(b);
    __builtin_memcpy(&k, &lo, 4);
  } else {
    __builtin_memcpy(&k, &b, 8);
  }
  return k;
}

__device__ __forceinline__ unsigned long long cmix(unsigned long long h) {
  h ^= h >> 33;
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  h *= 0xc4ceb9fe1a85ec53ull;
  h ^= h >> 33;
  return h;
}

// the last row in index order to hash to a slot owns it (:1197-1200)
template <class K>
__global__ void owner_kernel(const K* __restrict__ keys, size_t n, unsigned long long ts,
                             fhip_ht_slot* __restrict__ slots) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    atomicMax(&slots[home_slot(keys[i], ts)].owner, unsigned(i + 1));
}

// owners write their key and value into their slot
template <class K>
__global__ void fill_kernel(const K* __restrict__ keys, const uint64_t* __restrict__ vals,
                            size_t n, unsigned long long ts, fhip_ht_slot* __restrict__ slots) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const K k = keys[i];
    fhip_ht_slot* sl = &slots[home_slot(k, ts)];
    if (sl->owner == unsigned(i + 1)) {
      sl->key = key_to_bits(k);
      sl->val = vals ? vals[i] : 0;
    }
  }
}

// a non-owner whose key differs from its slot's goes to the conflict list;
// one equal to it is a duplicate (find_miss_unique_check, :302-433)
template <class K>
__global__ void classify_kernel(const K* __restrict__ keys, size_t n, unsigned long long ts,
                                const fhip_ht_slot* __restrict__ slots,
                                uint8_t* __restrict__ conflict, int* __restrict__ not_unique) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const K k = keys[i];
    const fhip_ht_slot* sl = &slots[home_slot(k, ts)];
    uint8_t c = 0;
    if (sl->owner != unsigned(i + 1)) {
      if (bits_to_key<K>(sl->key) == k) *not_unique = 1;
      else c = 1;
    }
    conflict[i] = c;
  }
}

struct FlagPred {
  const uint8_t* f;
  __device__ bool operator()(size_t i) const { return f[i] != 0; }
};

template <class K>
struct ConflictWriter {
  const K* keys;
  const uint64_t* vals;
  K* ck;
  uint64_t* cv;
  __device__ void operator()(size_t pos, size_t i) const {
    ck[pos] = keys[i];
    cv[pos] = vals ? vals[i] : 0;
  }
};

template <class K>
__global__ void adjacent_dup_kernel(const K* __restrict__ k, size_t n, int* __restrict__ flag) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x + 1; i < n; i += stride)
    if (!(k[i] != k[i - 1])) *flag = 1;
}

// secondary index: the first row of every run of equal keys in the sorted
// conflict list; the keys inserted are distinct, so a slot is claimed by
// its tag alone and readers only run after this kernel
template <class K>
__global__ void cidx_kernel(const K* __restrict__ ck, const uint64_t* __restrict__ cv, size_t m,
                            fhip_ht_cslot* __restrict__ tab, unsigned long long mask) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += stride) {
    const K k = ck[i];
    if (i > 0 && ck[i - 1] == k) continue;
    const unsigned long long kb = key_to_bits(k);
    unsigned long long s = cmix(kb) & mask;
    while (atomicCAS(&tab[s].tag, 0ull, (unsigned long long)(i + 1)) != 0ull) s = (s + 1) & mask;
    tab[s].key = kb;
    tab[s].val = cv ? cv[i] : 0;
  }
}

// lower_bound exactly as core/lower_bound.hpp:13-101 (`x <= s[mid]` goes left)
template <class K>
__device__ __forceinline__ size_t lb_search(const K* s, size_t m, K x) {
  size_t lo = 0, hi = m;
  while (lo < hi) {
    size_t mid = (lo + hi) >> 1;
    if (x <= s[mid]) hi = mid; else lo = mid + 1;
  }
  return lo;
}

template <class K>
__global__ __launch_bounds__(256) void probe_kernel(
    const K* __restrict__ q, size_t m, unsigned long long ts,
    const fhip_ht_slot* __restrict__ slots, const K* __restrict__ ck,
    const uint64_t* __restrict__ cv, size_t nconf, const fhip_ht_cslot* __restrict__ cidx,
    unsigned long long cmask, uint64_t* __restrict__ out, uint8_t* __restrict__ found) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += stride) {
    const K k = q[i];
    const fhip_ht_slot sl = slots[home_slot(k, ts)];
    uint64_t v = 0;
    uint8_t f = 0;
    if (sl.owner) {
      if (bits_to_key<K>(sl.key) == k) {
        v = sl.val;
        f = 1;
      } else if (nconf) {
        if (cidx) {
          const unsigned long long kb = key_to_bits(k);
          unsigned long long s = cmix(kb) & cmask;
          for (;;) {
            const fhip_ht_cslot c = cidx[s];
            if (c.tag == 0) break;
            if (c.key == kb) {
              v = c.val;
              f = 1;
              break;
            }
            s = (s + 1) & cmask;
          }
        } else {
          const size_t p = lb_search(ck, nconf, k);
          if (p < nconf && ck[p] == k) {
            v = cv ? cv[p] : 0;
            f = 1;
          }
        }
      }
    }
    if (out) out[i] = v;
    found[i] = f;
  }
}

__global__ void widen_kernel(const uint8_t* __restrict__ f, size_t n, int* __restrict__ out) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = f[i];
}

struct NotFound {
  const uint8_t* f;
  __device__ bool operator()(size_t i) const { return f[i] == 0; }
};

struct SlotOwned {
  const fhip_ht_slot* slots;
  __device__ bool operator()(size_t s) const { return slots[s].owner != 0; }
};

template <class K>
struct SlotKeyWriter {
  const fhip_ht_slot* slots;
  K* out;
  __device__ void operator()(size_t pos, size_t s) const { out[pos] = bits_to_key<K>(slots[s].key); }
};

template <class T>
T* dev_alloc(size_t count) {
  void* p = nullptr;
  if (count == 0) return nullptr;
  hipError_t e = hipMalloc(&p, count * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(FHIP_ERR_NOMEM, "unique_hashtable: device allocation failed");
  }
  return static_cast<T*>(p);
}

template <class K>
void build_t(fhip_ht* h, const K* keys, const uint64_t* vals, size_t n, bool check_unique,
             int* is_unique_ok, hipStream_t st) {
  const unsigned long long ts = h->table_size;
  h->slots = dev_alloc<fhip_ht_slot>(ts);
  FHIP_CHECK(hipMemsetAsync(h->slots, 0, ts * sizeof(fhip_ht_slot), st));
  DBuf flag(2 * sizeof(int)), conflict(std::max<size_t>(n, 1));
  FHIP_CHECK(hipMemsetAsync(flag.get(), 0, 2 * sizeof(int), st));
  if (n) {
    ProfScope prof("ht_build", st);
    owner_kernel<K><<<grid_for_n(n), 256, 0, st>>>(keys, n, ts, h->slots);
    FHIP_CHECK(hipGetLastError());
    fill_kernel<K><<<grid_for_n(n), 256, 0, st>>>(keys, vals, n, ts, h->slots);
    FHIP_CHECK(hipGetLastError());
    classify_kernel<K><<<grid_for_n(n), 256, 0, st>>>(keys, n, ts, h->slots,
                                                      conflict.as<uint8_t>(), flag.as<int>());
    FHIP_CHECK(hipGetLastError());
  }
  // conflict list: count, allocate, write in row order
  FlagPred pred{conflict.as<uint8_t>()};
  size_t nconf = n ? compact_apply(pred, ConflictWriter<K>{keys, vals, nullptr, nullptr}, n, 0, st) : 0;
  h->nconf = nconf;
  if (nconf) {
    h->conf_key = dev_alloc<K>(nconf);
    h->conf_val = dev_alloc<uint64_t>(nconf);
    compact_apply(pred, ConflictWriter<K>{keys, vals, (K*)h->conf_key, h->conf_val}, n, nconf, st);
    // radix_sort(conflict_key, conflict_val) (:1234)
    int kt = h->ktype;
    sort_pairs(kt, h->conf_key, h->conf_key, 8, h->conf_val, h->conf_val, nconf, false, false, st);
    if (check_unique && nconf > 1) {
      adjacent_dup_kernel<K><<<grid_for_n(nconf), 256, 0, st>>>((const K*)h->conf_key, nconf,
                                                                flag.as<int>());
      FHIP_CHECK(hipGetLastError());
    }
    if constexpr (std::is_integral<K>::value) {
      // floating keys keep the binary search: +0.0 and -0.0 compare equal
      // with different bits, and the reference's sort may place them apart
      size_t cap = 64;
      while (cap < 2 * nconf) cap <<= 1;
      h->cmask = cap - 1;
      h->cidx = dev_alloc<fhip_ht_cslot>(cap);
      FHIP_CHECK(hipMemsetAsync(h->cidx, 0, cap * sizeof(fhip_ht_cslot), st));
      cidx_kernel<K><<<grid_for_n(nconf), 256, 0, st>>>((const K*)h->conf_key, h->conf_val, nconf,
                                                        h->cidx, h->cmask);
      FHIP_CHECK(hipGetLastError());
    }
  }
  int dup = 0;
  FHIP_CHECK(hipMemcpyAsync(&dup, flag.get(), sizeof(int), hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  if (is_unique_ok) *is_unique_ok = dup ? 0 : 1;
}

}  // namespace

size_t table_size_for(size_t n) {
  const unsigned long long target = (unsigned long long)n * 3;  // HASH_TABLE_SIZE_MULT
  for (unsigned long long p : kPrimes)
    if (p >= target) return size_t(p);
  throw Error(FHIP_ERR_TOO_LARGE, "unique_hash_table: requested size too large");
}

fhip_ht* ht_build(int ktype, const void* keys, const uint64_t* vals, size_t n, bool check_unique,
                  int* is_unique_ok, hipStream_t st) {
  type_bytes(ktype);
  FHIP_REQUIRE(n < 0xFFFFFFFFull, FHIP_ERR_TOO_LARGE, "unique_hashtable: more than 2^32-2 rows");
  fhip_ht* h = new fhip_ht();
  h->ktype = ktype;
  h->has_vals = 1;
  h->n = n;
  try {
    h->table_size = table_size_for(n);
    dispatch_type(ktype, [&](auto z) {
      using K = decltype(z);
      build_t<K>(h, (const K*)keys, vals, n, check_unique, is_unique_ok, st);
    });
  } catch (...) {
    ht_destroy(h);
    throw;
  }
  return h;
}

void ht_destroy(fhip_ht* h) {
  if (!h) return;
  (void)hipFree(h->slots);
  (void)hipFree(h->conf_key);
  (void)hipFree(h->conf_val);
  (void)hipFree(h->cidx);
  delete h;
}

void ht_probe(const fhip_ht* h, const void* q, size_t m, uint64_t* vals_out, uint8_t* found,
              hipStream_t st) {
  if (m == 0) return;
  ProfScope prof("ht_probe", st);
  dispatch_type(h->ktype, [&](auto z) {
    using K = decltype(z);
    probe_kernel<K><<<grid_for_n(m), 256, 0, st>>>((const K*)q, m, h->table_size, h->slots,
                                                   (const K*)h->conf_key, h->conf_val, h->nconf,
                                                   h->cidx, h->cmask, vals_out, found);
  });
  FHIP_CHECK(hipGetLastError());
}

}  // namespace fhip

extern "C" {

int fhip_dev_ht_build(int ktype, const void* keys, const uint64_t* vals, size_t n,
                      int check_unique, int* is_unique_ok, fhip_ht** ht, void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(ht && (n == 0 || keys), FHIP_ERR_ARG, "ht_build: null pointer");
    *ht = fhip::ht_build(ktype, keys, vals, n, check_unique != 0, is_unique_ok,
                         fhip::as_stream(stream));
  });
}

int fhip_ht_destroy(fhip_ht* ht) {
  return fhip::guarded([&] { fhip::ht_destroy(ht); });
}

int fhip_ht_info(const fhip_ht* ht, size_t* table_size, size_t* conflict_size) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(ht, FHIP_ERR_ARG, "ht_info: null table");
    if (table_size) *table_size = ht->table_size;
    if (conflict_size) *conflict_size = ht->nconf;
  });
}

int fhip_dev_ht_lookup(const fhip_ht* ht, const void* keys, size_t n, uint64_t* vals_out,
                       size_t* miss_out, size_t* miss_n, void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(ht && vals_out && (n == 0 || keys), FHIP_ERR_ARG, "ht_lookup: null pointer");
    hipStream_t st = fhip::as_stream(stream);
    if (miss_n) *miss_n = 0;
    if (n == 0) return;
    fhip::DBuf fb(n);
    fhip::ht_probe(ht, keys, n, vals_out, fb.as<uint8_t>(), st);
    if (miss_out && miss_n)
      *miss_n = fhip::compact_indices(fhip::NotFound{fb.as<uint8_t>()}, n, miss_out, st);
  });
}

int fhip_dev_ht_check_existence(const fhip_ht* ht, const void* keys, size_t n, int* out,
                                void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(ht && out && (n == 0 || keys), FHIP_ERR_ARG, "ht_check_existence: null pointer");
    hipStream_t st = fhip::as_stream(stream);
    if (n == 0) return;
    fhip::DBuf fb(n);
    fhip::ht_probe(ht, keys, n, nullptr, fb.as<uint8_t>(), st);
    // widen the flags to int (std::vector<int> in the reference, :1517-1582)
    fhip::widen_kernel<<<fhip::grid_for_n(n), 256, 0, st>>>(fb.as<uint8_t>(), n, out);
    FHIP_CHECK(hipGetLastError());
  });
}

int fhip_dev_ht_all_keys(const fhip_ht* ht, void* out, size_t* out_n, void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(ht && out_n, FHIP_ERR_ARG, "ht_all_keys: null pointer");
    hipStream_t st = fhip::as_stream(stream);
    const int kb = fhip::type_bytes(ht->ktype);
    size_t owned = 0;
    fhip::dispatch_type(ht->ktype, [&](auto z) {
      using K = decltype(z);
      owned = fhip::compact_apply(fhip::SlotOwned{ht->slots},
                                  fhip::SlotKeyWriter<K>{ht->slots, (K*)out},
                                  ht->table_size, ht->n, st);
    });
    if (ht->nconf)
      FHIP_CHECK(hipMemcpyAsync(static_cast<char*>(out) + owned * kb, ht->conf_key,
                                ht->nconf * kb, hipMemcpyDeviceToDevice, st));
    FHIP_CHECK(hipStreamSynchronize(st));
    *out_n = owned + ht->nconf;
  });
}

}  // extern "C"
