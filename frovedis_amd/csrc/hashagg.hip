// Fused one-pass group_by + aggregation (hash path) for gfx950.
//
// Same result as dftable::group_by(k).select({k}, {sum_as(v), count_as(v),
// avg_as(v), max_as(i), ...}) on one rank (dftable.cc:346-366,3724-3772 ->
// dfcolumn_groupby.cc:16-120 -> dfcolumn_grouped_agg.cc): groups ascending by
// key, aggregates in argument order.  Instead of the reference's sort of
// (key, row) plus gathers, every row is read once:
//   1. each workgroup walks a contiguous row range and aggregates into an
//      LDS open-addressing table (heavy keys of skewed data stay on chip);
//      rows whose key finds no LDS slot go straight to the global table;
//   2. at the end the workgroup flushes its LDS entries into one global
//      open-addressing table in HBM (64-bit atomicCAS on the key, hardware
//      FP64 / integer atomics on the accumulators);
//   3. occupied entries are compacted, radix-sorted by key and gathered.
// Integer sums and counts are exact; FP64 sums use FP64 atomics (the order
// of additions is not fixed, so they match the reference fold to rounding).
#include "fhip_device.hpp"
#include "radix_sort.hpp"

#include <limits>

namespace fhip {
namespace {

constexpr int HA_NT = 512;
constexpr int HA_UNROLL = 4;
constexpr int HA_LDS_BITS = 11;
constexpr int HA_LDS_SLOTS = 1 << HA_LDS_BITS;
constexpr int HA_LDS_PROBES = 8;
constexpr int HA_MAX_ACC = 4;
constexpr unsigned long long HA_EMPTY = 0x8000000000000000ull;

enum { OP_SUM = 0, OP_COUNT = 1, OP_AVG = 2, OP_MAX = 3, OP_MIN = 4 };
enum { K_SUM = 0, K_MAX = 1, K_MIN = 2 };

struct AccDesc {
  const void* col;
  int kind;   // K_SUM / K_MAX / K_MIN
  int vtype;  // FHIP_* code
};

struct AccSet {
  AccDesc a[HA_MAX_ACC];
  int n;
};

__device__ __forceinline__ unsigned long long mix64(unsigned long long h) {
  h ^= h >> 33;
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  h *= 0xc4ceb9fe1a85ec53ull;
  h ^= h >> 33;
  return h;
}

template <class K>
__device__ __forceinline__ unsigned long long key_bits(K k) {
  // sign-extended for signed keys so that equal values give equal bits
  if constexpr (std::is_signed<K>::value) return (unsigned long long)(long long)k;
  else return (unsigned long long)k;
}

// accumulator identity as 64-bit pattern
__host__ __device__ inline unsigned long long acc_identity(int kind, int vtype) {
  if (kind == K_SUM) return 0ull;  // 0 for ints, +0.0 for FP
  const bool mx = kind == K_MAX;
  switch (vtype) {
    case FHIP_I32: { int v = mx ? std::numeric_limits<int>::lowest() : std::numeric_limits<int>::max(); return (unsigned long long)(unsigned)v; }
    case FHIP_U32: { unsigned v = mx ? 0u : 0xFFFFFFFFu; return v; }
    case FHIP_I64: { long long v = mx ? std::numeric_limits<long long>::lowest() : std::numeric_limits<long long>::max(); return (unsigned long long)v; }
    case FHIP_U64: return mx ? 0ull : ~0ull;
    case FHIP_F32: { float v = mx ? -std::numeric_limits<float>::max() : std::numeric_limits<float>::max(); unsigned b; std::memcpy(&b, &v, 4); return b; }
    default: { double v = mx ? -std::numeric_limits<double>::max() : std::numeric_limits<double>::max(); unsigned long long b; std::memcpy(&b, &v, 8); return b; }
  }
}

// Loads one row of a value column as a 64-bit pattern.
__device__ __forceinline__ unsigned long long load_bits(const void* col, int vtype, size_t row) {
  if (vtype == FHIP_I32 || vtype == FHIP_U32 || vtype == FHIP_F32)
    return static_cast<const unsigned*>(col)[row];
  return static_cast<const unsigned long long*>(col)[row];
}

// Merge an accumulator word (already aggregated) into another.
__device__ __forceinline__ void acc_merge(unsigned long long* w, int kind, int vtype,
                                          unsigned long long src) {
  switch (vtype) {
    case FHIP_F64: {
      double v;
      __builtin_memcpy(&v, &src, 8);
      double* p = reinterpret_cast<double*>(w);
      if (kind == K_SUM) unsafeAtomicAdd(p, v);
      else if (v != v) return;  // NaN is never selected by `if (m < v)`
      else if (kind == K_MAX) atomicMax(p, v);
      else atomicMin(p, v);
      return;
    }
    case FHIP_F32: {
      float v;
      unsigned b = unsigned(src);
      __builtin_memcpy(&v, &b, 4);
      float* p = reinterpret_cast<float*>(w);
      if (kind == K_SUM) unsafeAtomicAdd(p, v);
      else if (v != v) return;
      else if (kind == K_MAX) atomicMax(p, v);
      else atomicMin(p, v);
      return;
    }
    case FHIP_I32: {
      int v = int(unsigned(src));
      int* p = reinterpret_cast<int*>(w);
      if (kind == K_SUM) atomicAdd(p, v);
      else if (kind == K_MAX) atomicMax(p, v);
      else atomicMin(p, v);
      return;
    }
    case FHIP_U32: {
      unsigned v = unsigned(src);
      unsigned* p = reinterpret_cast<unsigned*>(w);
      if (kind == K_SUM) atomicAdd(p, v);
      else if (kind == K_MAX) atomicMax(p, v);
      else atomicMin(p, v);
      return;
    }
    case FHIP_I64: {
      long long v = (long long)src;
      long long* p = reinterpret_cast<long long*>(w);
      if (kind == K_SUM) atomicAdd(reinterpret_cast<unsigned long long*>(p), src);
      else if (kind == K_MAX) atomicMax(p, v);
      else atomicMin(p, v);
      return;
    }
    default: {
      if (kind == K_SUM) atomicAdd(w, src);
      else if (kind == K_MAX) atomicMax(w, src);
      else atomicMin(w, src);
      return;
    }
  }
}

// Global table: `stride` 64-bit words per entry: key, count, acc[0..n).
struct GTable {
  unsigned long long* w;
  unsigned long long mask;  // capacity - 1; entry `capacity` is the EMPTY-key entry
  int stride;
  unsigned long long* inserted;
  int* overflow;
};

constexpr unsigned long long HA_MAX_PROBES = 4096;

// Finds or inserts the entry of key bits `kb`; -1 when the table is too full
// (the caller retries with a larger table).
__device__ __forceinline__ long long gt_find(const GTable& t, unsigned long long kb,
                                             unsigned long long h) {
  if (kb == HA_EMPTY) return (long long)(t.mask + 1);
  unsigned long long i = h & t.mask;
  const unsigned long long limit = t.mask < HA_MAX_PROBES ? t.mask + 1 : HA_MAX_PROBES;
  for (unsigned long long p = 0; p < limit; p++) {
    unsigned long long* slot = t.w + i * t.stride;
    unsigned long long cur = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == kb) return (long long)i;
    if (cur == HA_EMPTY) {
      unsigned long long old = atomicCAS(slot, HA_EMPTY, kb);
      if (old == HA_EMPTY) {
        unsigned long long c = atomicAdd(t.inserted, 1ull);
        if (c + 1 > ((t.mask + 1) / 4) * 3) atomicExch(t.overflow, 1);
        return (long long)i;
      }
      if (old == kb) return (long long)i;
    }
    i = (i + 1) & t.mask;
  }
  atomicExch(t.overflow, 1);
  return -1;
}

__device__ __forceinline__ bool gt_overflowed(const GTable& t) {
  return __hip_atomic_load(t.overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

template <class K>
__global__ __launch_bounds__(HA_NT) void hashagg_kernel(const K* __restrict__ keys, size_t n,
                                                        size_t rows_per_block, AccSet accs,
                                                        GTable gt) {
  // LDS table: keys, per-slot row counts, then na accumulator arrays
  extern __shared__ unsigned long long lds[];
  unsigned long long* lkey = lds;
  unsigned long long* lacc = lds + HA_LDS_SLOTS;  // [na][slots]
  const int na = accs.n;
  unsigned* lcnt = reinterpret_cast<unsigned*>(lacc + size_t(na) * HA_LDS_SLOTS);
  for (int s = threadIdx.x; s < HA_LDS_SLOTS; s += HA_NT) {
    lkey[s] = HA_EMPTY;
    lcnt[s] = 0;
    for (int a = 0; a < na; a++)
      lacc[a * HA_LDS_SLOTS + s] = acc_identity(accs.a[a].kind, accs.a[a].vtype);
  }
  __syncthreads();

  const size_t beg = size_t(blockIdx.x) * rows_per_block;
  const size_t end = beg + rows_per_block < n ? beg + rows_per_block : n;
  for (size_t r0 = beg + threadIdx.x; r0 < end; r0 += size_t(HA_NT) * HA_UNROLL) {
    if (gt_overflowed(gt)) break;  // result will be discarded and rebuilt
    unsigned long long kbv[HA_UNROLL];
    unsigned long long vv[HA_UNROLL][HA_MAX_ACC];
#pragma unroll
    for (int u = 0; u < HA_UNROLL; u++) {
      const size_t r = r0 + size_t(u) * HA_NT;
      kbv[u] = r < end ? key_bits(keys[r]) : 0ull;
#pragma unroll
      for (int a = 0; a < HA_MAX_ACC; a++)
        vv[u][a] = (a < na && r < end) ? load_bits(accs.a[a].col, accs.a[a].vtype, r) : 0ull;
    }
#pragma unroll
    for (int u = 0; u < HA_UNROLL; u++) {
      const size_t r = r0 + size_t(u) * HA_NT;
      if (r >= end) break;
      const unsigned long long kb = kbv[u];
      const unsigned long long h = mix64(kb);
      int slot = -1;
      if (kb != HA_EMPTY) {
        int s = int(h >> (64 - HA_LDS_BITS));
#pragma unroll 1
        for (int p = 0; p < HA_LDS_PROBES; p++) {
          unsigned long long cur = lkey[s];
          if (cur == kb) { slot = s; break; }
          if (cur == HA_EMPTY) {
            unsigned long long old = atomicCAS(&lkey[s], HA_EMPTY, kb);
            if (old == HA_EMPTY || old == kb) { slot = s; break; }
          }
          s = (s + 1) & (HA_LDS_SLOTS - 1);
        }
      }
      if (slot >= 0) {
        atomicAdd(&lcnt[slot], 1u);
        for (int a = 0; a < na; a++)
          acc_merge(&lacc[a * HA_LDS_SLOTS + slot], accs.a[a].kind, accs.a[a].vtype, vv[u][a]);
      } else {
        long long e = gt_find(gt, kb, h);
        if (e < 0) continue;
        unsigned long long* ent = gt.w + size_t(e) * gt.stride;
        atomicAdd(ent + 1, 1ull);
        for (int a = 0; a < na; a++)
          acc_merge(ent + 2 + a, accs.a[a].kind, accs.a[a].vtype, vv[u][a]);
      }
    }
  }
  __syncthreads();
  for (int s = threadIdx.x; s < HA_LDS_SLOTS; s += HA_NT) {
    const unsigned long long kb = lkey[s];
    if (kb == HA_EMPTY || gt_overflowed(gt)) continue;
    long long e = gt_find(gt, kb, mix64(kb));
    if (e < 0) continue;
    unsigned long long* ent = gt.w + size_t(e) * gt.stride;
    atomicAdd(ent + 1, (unsigned long long)lcnt[s]);
    for (int a = 0; a < na; a++)
      acc_merge(ent + 2 + a, accs.a[a].kind, accs.a[a].vtype, lacc[a * HA_LDS_SLOTS + s]);
  }
}

__global__ void gt_init_kernel(unsigned long long* w, size_t entries, int stride, AccSet accs) {
  const size_t stride_g = size_t(gridDim.x) * blockDim.x;
  for (size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x; e < entries; e += stride_g) {
    unsigned long long* ent = w + e * stride;
    ent[0] = HA_EMPTY;
    ent[1] = 0;
    for (int a = 0; a < accs.n; a++) ent[2 + a] = acc_identity(accs.a[a].kind, accs.a[a].vtype);
  }
}

struct Occupied {
  const unsigned long long* w;
  int stride;
  size_t cap;
  __device__ bool operator()(size_t e) const {
    // the EMPTY-key entry (index cap) is occupied when its count is non-zero
    return e == cap ? w[e * stride + 1] != 0 : w[e * stride] != HA_EMPTY;
  }
};

template <class K>
__global__ void gather_group_keys(const unsigned long long* __restrict__ w, int stride,
                                  const size_t* __restrict__ ent, size_t G, K* __restrict__ keys) {
  const size_t st = size_t(gridDim.x) * blockDim.x;
  for (size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x; g < G; g += st)
    keys[g] = K(w[ent[g] * stride]);
}

// Writes one aggregate in sorted-group order.
__global__ void emit_kernel(const unsigned long long* __restrict__ w, int stride,
                            const size_t* __restrict__ ent, size_t G, int op, int vtype, int slot,
                            void* __restrict__ out) {
  const size_t st = size_t(gridDim.x) * blockDim.x;
  for (size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x; g < G; g += st) {
    const unsigned long long* e = w + ent[g] * stride;
    const unsigned long long cnt = e[1];
    if (op == OP_COUNT) {
      static_cast<unsigned long long*>(out)[g] = cnt;
      continue;
    }
    const unsigned long long a = e[2 + slot];
    if (op == OP_AVG) {
      double s;
      switch (vtype) {
        case FHIP_I32: s = double(int(unsigned(a))); break;
        case FHIP_U32: s = double(unsigned(a)); break;
        case FHIP_I64: s = double((long long)a); break;
        case FHIP_U64: s = double(a); break;
        case FHIP_F32: { float f; unsigned b = unsigned(a); __builtin_memcpy(&f, &b, 4); s = double(f); break; }
        default: __builtin_memcpy(&s, &a, 8); break;
      }
      static_cast<double*>(out)[g] = s / double(cnt);
      continue;
    }
    switch (vtype) {
      case FHIP_I32: case FHIP_U32: static_cast<unsigned*>(out)[g] = unsigned(a); break;
      case FHIP_F32: {
        float f; unsigned b = unsigned(a); __builtin_memcpy(&f, &b, 4);
        if (op == OP_SUM) f = 0.0f + f;
        static_cast<float*>(out)[g] = f;
        break;
      }
      case FHIP_F64: {
        double d; __builtin_memcpy(&d, &a, 8);
        if (op == OP_SUM) d = 0.0 + d;
        static_cast<double*>(out)[g] = d;
        break;
      }
      default: static_cast<unsigned long long*>(out)[g] = a; break;
    }
  }
}

int grid_for(size_t n, int per_cu = 8) {
  return int(std::max<size_t>(1, std::min<size_t>(ceil_div(n, 256), size_t(fill_blocks(256, per_cu)))));
}

}  // namespace

// Sorted-by-key result of a fused group-by; see fhip_dev_group_by_agg.
void group_by_agg(int ktype, const void* keys, size_t n, const fhip_agg* aggs, int naggs,
                  void* keys_out, size_t capacity, size_t* ngroups, size_t table_hint,
                  hipStream_t st) {
  FHIP_REQUIRE(ktype == FHIP_I32 || ktype == FHIP_U32 || ktype == FHIP_I64 || ktype == FHIP_U64,
               FHIP_ERR_TYPE, "group_by_agg: integer keys only (use group_by for FP keys)");
  FHIP_REQUIRE(naggs >= 0 && naggs <= 16, FHIP_ERR_ARG, "group_by_agg: too many aggregations");
  // distinct accumulators: (kind, column); count and avg reuse them
  AccSet accs{};
  std::vector<int> slot_of(naggs, -1);
  for (int i = 0; i < naggs; i++) {
    const fhip_agg& a = aggs[i];
    FHIP_REQUIRE(a.op >= OP_SUM && a.op <= OP_MIN, FHIP_ERR_ARG, "group_by_agg: bad op");
    if (a.op == OP_COUNT) continue;
    type_bytes(a.vtype);
    FHIP_REQUIRE(a.col, FHIP_ERR_ARG, "group_by_agg: null column");
    int kind = (a.op == OP_MAX) ? K_MAX : (a.op == OP_MIN) ? K_MIN : K_SUM;
    int found = -1;
    for (int j = 0; j < accs.n; j++)
      if (accs.a[j].col == a.col && accs.a[j].kind == kind && accs.a[j].vtype == a.vtype) found = j;
    if (found < 0) {
      FHIP_REQUIRE(accs.n < HA_MAX_ACC, FHIP_ERR_ARG, "group_by_agg: more than 4 accumulators");
      accs.a[accs.n] = AccDesc{a.col, kind, a.vtype};
      found = accs.n++;
    }
    slot_of[i] = found;
  }
  *ngroups = 0;
  if (n == 0) return;
  const int stride = 2 + accs.n;
  size_t cap = 1024;
  size_t want = table_hint ? table_hint : std::min<size_t>(n, size_t(1) << 21);
  while (cap < 2 * want) cap <<= 1;
  const int grid = fill_blocks(HA_NT, 2);
  const size_t lds_bytes = size_t(HA_LDS_SLOTS) * (8 + 8 * size_t(accs.n) + 4);
  static bool lds_attr_set = false;
  if (!lds_attr_set) {  // allow more than 64 KiB of dynamic LDS (4 accumulators)
    const int max_lds = int(size_t(HA_LDS_SLOTS) * (8 + 8 * HA_MAX_ACC + 4));
    FHIP_CHECK(hipFuncSetAttribute((const void*)hashagg_kernel<int32_t>, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
    FHIP_CHECK(hipFuncSetAttribute((const void*)hashagg_kernel<uint32_t>, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
    FHIP_CHECK(hipFuncSetAttribute((const void*)hashagg_kernel<int64_t>, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
    FHIP_CHECK(hipFuncSetAttribute((const void*)hashagg_kernel<uint64_t>, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
    lds_attr_set = true;
  }
  const size_t rows_per_block = ceil_div(n, size_t(grid));
  DBuf tab, flags(2 * sizeof(unsigned long long));
  for (;;) {
    tab.reset((cap + 1) * stride * sizeof(unsigned long long));
    GTable gt{tab.as<unsigned long long>(), (unsigned long long)(cap - 1), stride,
              flags.as<unsigned long long>(), reinterpret_cast<int*>(flags.as<unsigned long long>() + 1)};
    FHIP_CHECK(hipMemsetAsync(flags.get(), 0, 2 * sizeof(unsigned long long), st));
    gt_init_kernel<<<grid_for(cap + 1), 256, 0, st>>>(gt.w, cap + 1, stride, accs);
    FHIP_CHECK(hipGetLastError());
    ProfScope prof("hashagg", st);
    switch (ktype) {
      case FHIP_I32: hashagg_kernel<int32_t><<<grid, HA_NT, lds_bytes, st>>>((const int32_t*)keys, n, rows_per_block, accs, gt); break;
      case FHIP_U32: hashagg_kernel<uint32_t><<<grid, HA_NT, lds_bytes, st>>>((const uint32_t*)keys, n, rows_per_block, accs, gt); break;
      case FHIP_I64: hashagg_kernel<int64_t><<<grid, HA_NT, lds_bytes, st>>>((const int64_t*)keys, n, rows_per_block, accs, gt); break;
      default: hashagg_kernel<uint64_t><<<grid, HA_NT, lds_bytes, st>>>((const uint64_t*)keys, n, rows_per_block, accs, gt); break;
    }
    FHIP_CHECK(hipGetLastError());
    int ovf = 0;
    FHIP_CHECK(hipMemcpyAsync(&ovf, gt.overflow, sizeof(int), hipMemcpyDeviceToHost, st));
    FHIP_CHECK(hipStreamSynchronize(st));
    if (!ovf) break;
    FHIP_REQUIRE(cap < (size_t(1) << 40), FHIP_ERR_TOO_LARGE, "group_by_agg: table too large");
    cap <<= 2;  // too many groups for this table: retry 4x larger
  }
  // compact occupied entries, then sort them by key
  const size_t entries = cap + 1;
  DBuf entb(entries * sizeof(size_t));
  size_t G = compact_indices(Occupied{tab.as<unsigned long long>(), stride, cap}, entries,
                             entb.as<size_t>(), st);
  *ngroups = G;  // reported even when it does not fit, so the caller can resize
  FHIP_REQUIRE(G <= capacity, FHIP_ERR_TOO_LARGE,
               "group_by_agg: output capacity " + std::to_string(capacity) + " < groups " +
                   std::to_string(G));
  if (G == 0) return;
  const int kb = type_bytes(ktype);
  DBuf gk(G * size_t(kb));
  switch (ktype) {
    case FHIP_I32: gather_group_keys<int32_t><<<grid_for(G), 256, 0, st>>>(tab.as<unsigned long long>(), stride, entb.as<size_t>(), G, gk.as<int32_t>()); break;
    case FHIP_U32: gather_group_keys<uint32_t><<<grid_for(G), 256, 0, st>>>(tab.as<unsigned long long>(), stride, entb.as<size_t>(), G, gk.as<uint32_t>()); break;
    case FHIP_I64: gather_group_keys<int64_t><<<grid_for(G), 256, 0, st>>>(tab.as<unsigned long long>(), stride, entb.as<size_t>(), G, gk.as<int64_t>()); break;
    default: gather_group_keys<uint64_t><<<grid_for(G), 256, 0, st>>>(tab.as<unsigned long long>(), stride, entb.as<size_t>(), G, gk.as<uint64_t>()); break;
  }
  FHIP_CHECK(hipGetLastError());
  // keys are distinct, so any stable order is the reference's ascending order
  sort_pairs(ktype, gk.get(), gk.get(), 8, entb.get(), entb.get(), G, false, false, st);
  FHIP_CHECK(hipMemcpyAsync(keys_out, gk.get(), G * size_t(kb), hipMemcpyDeviceToDevice, st));
  for (int i = 0; i < naggs; i++) {
    emit_kernel<<<grid_for(G), 256, 0, st>>>(tab.as<unsigned long long>(), stride,
                                             entb.as<size_t>(), G, aggs[i].op, aggs[i].vtype,
                                             slot_of[i] < 0 ? 0 : slot_of[i], aggs[i].out);
    FHIP_CHECK(hipGetLastError());
  }
}

}  // namespace fhip

extern "C" int fhip_dev_group_by_agg(int ktype, const void* keys, size_t n, const fhip_agg* aggs,
                                     int naggs, void* keys_out, size_t capacity,
                                     size_t* ngroups, size_t table_hint, void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(ngroups && (naggs == 0 || aggs), FHIP_ERR_ARG, "group_by_agg: null pointer");
    FHIP_REQUIRE(n == 0 || (keys && keys_out), FHIP_ERR_ARG, "group_by_agg: null keys");
    fhip::group_by_agg(ktype, keys, n, aggs, naggs, keys_out, capacity, ngroups, table_hint,
                       fhip::as_stream(stream));
  });
}
