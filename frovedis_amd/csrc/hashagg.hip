// Fused group_by + aggregation (hash path) for gfx950.
//
// Same result as dftable::group_by(k).select({k}, {sum_as(v), count_as(v),
// avg_as(v), max_as(i), ...}) on one rank (dftable.cc:346-366,3724-3772 ->
// dfcolumn_groupby.cc:16-120 -> dfcolumn_grouped_agg.cc): groups ascending by
// key, aggregates in argument order.  Instead of the reference's sort of
// (key, row) plus gathers, the rows are hash-partitioned once so that every
// partition's groups fit one workgroup's LDS table:
//   1. pa_count: per 4096-row tile, rows per partition (P = 2^pbits
//      partitions by the top bits of a 64-bit mix of the key);
//   2. prefix_sum over the (partition, tile) counts; pa_scatter: each tile
//      stages its key and value columns in LDS in partition order and writes
//      them out as contiguous runs (coalesced);
//   3. pa_agg: one workgroup per chunk (at most `groupby.chunk` rows of one
//      partition) aggregates into an LDS open-addressing table; a wave whose
//      rows all hit one slot (a heavy key) reduces in registers first;
//   4. pa_merge: one workgroup per partition merges its chunk partials in
//      LDS; then the groups are compacted, their keys radix-sorted and the
//      aggregates gathered in key order.
// A partition with more groups than its table raises a flag; the pipeline
// reruns with 4x the partitions, and past 4096 partitions falls back to the
// sort path.  Counts, integer sums and max/min are exact; FP sums use LDS
// FP64 atomics (the order of additions is not fixed, so they match the
// reference fold to rounding).
#include "fhip_device.hpp"
#include "radix_sort.hpp"

#include <cstring>
#include <limits>

namespace fhip {
namespace {

constexpr int HA_MAX_ACC = 4;
constexpr int PA_NT = 1024;           // threads of the aggregate / merge kernels (one per CU: LDS)
constexpr int PA_UNROLL = 4;          // rows in flight per thread (aggregate)
constexpr int PS_NT = 1024;           // threads of the count / scatter kernels
constexpr int PS_ITEMS = 4;           // rows per thread of the count kernel
constexpr int PS_TILE = PS_NT * PS_ITEMS;  // rows per count tile; spans are whole count tiles
// rows per thread of the scatter kernel: 12 (12288-row tiles, longer runs
// per partition) while the partitions leave LDS for it, else 4
constexpr int PS_ITEMS_WIDE = 12;
constexpr int PS_WIDE_MAX_P = 2048;
constexpr int PS_SPAN = 16;            // default tiles per count / scatter workgroup (a "span")
constexpr unsigned long long PA_CHUNK = 1ull << 17;  // rows per aggregate workgroup
constexpr int PA_SLOTS = 2048;        // LDS table slots (power of two)
constexpr int PA_MAX_PBITS = 12;      // 4096 partitions (LDS of the scatter kernel)
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

__host__ __device__ __forceinline__ unsigned long long mix64(unsigned long long h) {
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

// acc_merge for U rows of one accumulator (slots base[idx[u]], rows with
// in[u]), the type dispatch done once for all of them.
template <int U>
__device__ __forceinline__ void acc_merge_rows(unsigned long long* base, const unsigned* idx,
                                               const bool* in, int kind, int vtype,
                                               const unsigned long long* v) {
#define FHIP_MERGE_ROWS(T, CONV, OP)                                  \
  {                                                                   \
    _Pragma("unroll") for (int u = 0; u < U; u++) if (in[u]) {        \
      T* p = reinterpret_cast<T*>(base + idx[u]);                     \
      const T x = CONV(v[u]);                                         \
      OP;                                                             \
    }                                                                 \
    return;                                                           \
  }
  auto as_i32 = [](unsigned long long b) { return int(unsigned(b)); };
  auto as_u32 = [](unsigned long long b) { return unsigned(b); };
  auto as_i64 = [](unsigned long long b) { return (long long)b; };
  auto as_u64 = [](unsigned long long b) { return b; };
  auto as_f32 = [](unsigned long long b) { float f; unsigned w = unsigned(b); __builtin_memcpy(&f, &w, 4); return f; };
  auto as_f64 = [](unsigned long long b) { double d; __builtin_memcpy(&d, &b, 8); return d; };
  switch (vtype) {
    case FHIP_F64:
      if (kind == K_SUM) FHIP_MERGE_ROWS(double, as_f64, unsafeAtomicAdd(p, x))
      else if (kind == K_MAX) FHIP_MERGE_ROWS(double, as_f64, if (x == x) atomicMax(p, x))
      else FHIP_MERGE_ROWS(double, as_f64, if (x == x) atomicMin(p, x))
    case FHIP_F32:
      if (kind == K_SUM) FHIP_MERGE_ROWS(float, as_f32, unsafeAtomicAdd(p, x))
      else if (kind == K_MAX) FHIP_MERGE_ROWS(float, as_f32, if (x == x) atomicMax(p, x))
      else FHIP_MERGE_ROWS(float, as_f32, if (x == x) atomicMin(p, x))
    case FHIP_I32:
      if (kind == K_SUM) FHIP_MERGE_ROWS(int, as_i32, atomicAdd(p, x))
      else if (kind == K_MAX) FHIP_MERGE_ROWS(int, as_i32, atomicMax(p, x))
      else FHIP_MERGE_ROWS(int, as_i32, atomicMin(p, x))
    case FHIP_U32:
      if (kind == K_SUM) FHIP_MERGE_ROWS(unsigned, as_u32, atomicAdd(p, x))
      else if (kind == K_MAX) FHIP_MERGE_ROWS(unsigned, as_u32, atomicMax(p, x))
      else FHIP_MERGE_ROWS(unsigned, as_u32, atomicMin(p, x))
    case FHIP_I64:
      if (kind == K_SUM) FHIP_MERGE_ROWS(unsigned long long, as_u64, atomicAdd(p, x))
      else if (kind == K_MAX) FHIP_MERGE_ROWS(long long, as_i64, atomicMax(p, x))
      else FHIP_MERGE_ROWS(long long, as_i64, atomicMin(p, x))
    default:
      if (kind == K_SUM) FHIP_MERGE_ROWS(unsigned long long, as_u64, atomicAdd(p, x))
      else if (kind == K_MAX) FHIP_MERGE_ROWS(unsigned long long, as_u64, atomicMax(p, x))
      else FHIP_MERGE_ROWS(unsigned long long, as_u64, atomicMin(p, x))
  }
#undef FHIP_MERGE_ROWS
}

// ---- pass 1: partition histogram -------------------------------------------
// partition of a key: the top pbits of its hash; LDS slots use the low bits
__device__ __forceinline__ unsigned part_of(unsigned long long h, int pbits) {
  return pbits ? unsigned(h >> (64 - pbits)) : 0u;
}

// counts per (partition, span), partition-major: cnt[p * spans + s]; a span
// is span_tiles consecutive tiles, counted by one workgroup
template <class K>
__global__ __launch_bounds__(PS_NT) void pa_count_kernel(const K* __restrict__ keys, size_t n,
                                                         int pbits, size_t spans, int span_tiles,
                                                         unsigned long long* __restrict__ cnt,
                                                         const unsigned long long* __restrict__ seg_n,
                                                         size_t seg_stride) {
  extern __shared__ unsigned lh[];
  const int P = 1 << pbits;
  for (int i = threadIdx.x; i < P; i += PS_NT) lh[i] = 0;
  __syncthreads();
  const unsigned lane = __lane_id();
  const size_t r0 = seg_n ? size_t(blockIdx.x) * seg_stride : size_t(blockIdx.x) * span_tiles * PS_TILE;
  const size_t r1 = seg_n ? r0 + seg_n[blockIdx.x] : std::min(n, r0 + size_t(span_tiles) * PS_TILE);
  for (size_t t0 = r0; t0 < r1; t0 += PS_TILE) {
#pragma unroll 4
    for (int j = 0; j < PS_ITEMS; j++) {
      const size_t i = t0 + size_t(j) * PS_NT + threadIdx.x;
      const bool ok = i < r1;
      const unsigned pp = ok ? part_of(mix64(key_bits(keys[i])), pbits) : 0u;
      const unsigned long long act = __ballot(ok);
      if (!act) continue;
      // the lanes in the first active lane's partition count with one atomic
      const int leader = __builtin_ffsll((long long)act) - 1;
      const unsigned p0 = __shfl(pp, leader, 64);
      const unsigned long long gm = __ballot(ok && pp == p0);
      if (lane == unsigned(leader)) atomicAdd(&lh[p0], unsigned(__popcll(gm)));
      else if (ok && pp != p0) atomicAdd(&lh[pp], 1u);
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < P; q += PS_NT) cnt[size_t(q) * spans + blockIdx.x] = lh[q];
}

// ---- pass 2: scatter rows (key + accumulator columns) into partitions ------
struct Cols {
  const void* in[HA_MAX_ACC];
  void* out[HA_MAX_ACC];
  int bytes[HA_MAX_ACC];
  int n;
};

// One span (span_tiles * PS_TILE rows, or a cold region) in tiles of
// PS_NT * ITEMS rows, tile by tile: rows are ranked
// by partition in LDS, staged in partition order and written out column by
// column in runs, so consecutive lanes store consecutive addresses.
// incl/cnt: inclusive scan and counts of the partition-major (partition,
// span) count matrix (exact offsets, no global atomics); the running offset
// of every partition advances by each tile's count.
template <class K, int ITEMS>
__global__ __launch_bounds__(PS_NT) void pa_scatter_kernel(
    const K* __restrict__ keys, size_t n, int pbits, size_t spans, int span_tiles, Cols cols,
    const unsigned long long* __restrict__ cnt, const unsigned long long* __restrict__ incl,
    K* __restrict__ out_keys, const unsigned long long* __restrict__ seg_n, size_t seg_stride) {
  extern __shared__ unsigned long long lds_u64[];
  constexpr int TILE = PS_NT * ITEMS;
  // value columns held in registers from the start of a tile (the others
  // are loaded when their turn to be staged comes)
  constexpr int PRELOAD = ITEMS <= 4 ? 2 : 0;
  const int P = 1 << pbits;
  unsigned long long* stage = lds_u64;                                     // [TILE]
  unsigned long long* gbase = stage + TILE;                                // [P]
  unsigned* pc = reinterpret_cast<unsigned*>(gbase + P);                   // [P]
  unsigned* ps = pc + P;                                                   // [P]
  unsigned short* pid = reinterpret_cast<unsigned short*>(ps + P);         // [TILE]
  __shared__ unsigned scan_tmp[PS_NT / 64];
  const unsigned tid = threadIdx.x, lane = __lane_id();
  const size_t sp = blockIdx.x;
  for (int q = tid; q < P; q += PS_NT) {
    const size_t m = size_t(q) * spans + sp;
    gbase[q] = incl[m] - cnt[m];  // exclusive offset of (partition q, this span)
    pc[q] = 0;
  }
  const size_t r0 = seg_n ? sp * seg_stride : sp * span_tiles * PS_TILE;
  const size_t r1 = seg_n ? r0 + seg_n[sp] : std::min(n, r0 + size_t(span_tiles) * PS_TILE);
  const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  for (size_t t0 = r0; t0 < r1; t0 += TILE) {
    __syncthreads();
    const unsigned tn = unsigned(r1 - t0 < size_t(TILE) ? r1 - t0 : size_t(TILE));
    unsigned part[ITEMS];
    unsigned short slot[ITEMS];
    // the key and the first PRELOAD columns of the tile are loaded together
    K kreg[ITEMS];
    unsigned long long creg[PRELOAD > 0 ? PRELOAD : 1][ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
      const unsigned r = unsigned(j) * PS_NT + tid;
      kreg[j] = r < tn ? keys[t0 + r] : K(0);
    }
#pragma unroll
    for (int a = 0; a < PRELOAD; a++) {
      if (a >= cols.n) break;
#pragma unroll
      for (int j = 0; j < ITEMS; j++) {
        const unsigned r = unsigned(j) * PS_NT + tid;
        creg[a][j] = r >= tn ? 0ull
                     : cols.bytes[a] == 4
                         ? (unsigned long long)static_cast<const unsigned*>(cols.in[a])[t0 + r]
                         : static_cast<const unsigned long long*>(cols.in[a])[t0 + r];
      }
    }
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
      const unsigned r = unsigned(j) * PS_NT + tid;
      part[j] = r < tn ? part_of(mix64(key_bits(kreg[j])), pbits) : 0u;
    }
    // counts of this tile per partition
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
      const unsigned r = unsigned(j) * PS_NT + tid;
      const bool ok = r < tn;
      const unsigned long long act = __ballot(ok);
      if (!act) continue;
      const int leader = __builtin_ffsll((long long)act) - 1;
      const unsigned p0 = __shfl(part[j], leader, 64);
      const unsigned long long gm = __ballot(ok && part[j] == p0);
      if (lane == unsigned(leader)) atomicAdd(&pc[p0], unsigned(__popcll(gm)));
      else if (ok && part[j] != p0) atomicAdd(&pc[part[j]], 1u);
    }
    __syncthreads();
    // tile-local start of each partition (block scan, P / PS_NT entries per thread)
    {
      const int per = (P + PS_NT - 1) / PS_NT;
      unsigned local = 0;
      for (int e = 0; e < per; e++) {
        const int q = tid * per + e;
        if (q < P) local += pc[q];
      }
      unsigned run = block_excl_scan<PS_NT>(local, scan_tmp, (unsigned*)nullptr);
      for (int e = 0; e < per; e++) {
        const int q = tid * per + e;
        if (q < P) {
          ps[q] = run;
          run += pc[q];
        }
      }
    }
    __syncthreads();
    for (int q = tid; q < P; q += PS_NT) pc[q] = 0;
    __syncthreads();
    // rank within (tile, partition)
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
      const unsigned r = unsigned(j) * PS_NT + tid;
      const bool ok = r < tn;
      const unsigned long long act = __ballot(ok);
      slot[j] = 0;
      if (!act) continue;
      // the lanes in the first active lane's partition reserve one block
      const int leader = __builtin_ffsll((long long)act) - 1;
      const unsigned p0 = __shfl(part[j], leader, 64);
      const bool in0 = ok && part[j] == p0;
      const unsigned long long gm = __ballot(in0);
      unsigned base = 0;
      if (lane == unsigned(leader)) base = atomicAdd(&pc[p0], unsigned(__popcll(gm)));
      base = __shfl(base, leader, 64);
      if (in0) slot[j] = (unsigned short)(ps[p0] + base + __popcll(gm & lt));
      else if (ok) slot[j] = (unsigned short)(ps[part[j]] + atomicAdd(&pc[part[j]], 1u));
    }
    // keys: stage in partition order, then write runs; every thread keeps
    // the destinations of its staged positions for all the columns
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
      const unsigned r = unsigned(j) * PS_NT + tid;
      if (r < tn) {
        stage[slot[j]] = (unsigned long long)kreg[j];
        pid[slot[j]] = (unsigned short)part[j];
      }
    }
    __syncthreads();
    unsigned long long dst[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
      const unsigned i = unsigned(k) * PS_NT + tid;
      dst[k] = 0;
      if (i < tn) {
        const unsigned q = pid[i];
        dst[k] = gbase[q] + (i - ps[q]);
        out_keys[dst[k]] = K(stage[i]);
      }
    }
    for (int a = 0; a < cols.n; a++) {
      __syncthreads();
#pragma unroll
      for (int j = 0; j < ITEMS; j++) {
        const unsigned r = unsigned(j) * PS_NT + tid;
        if (r < tn)
          stage[slot[j]] = a < PRELOAD ? creg[a < PRELOAD ? a : 0][j]
                           : cols.bytes[a] == 4
                               ? (unsigned long long)static_cast<const unsigned*>(cols.in[a])[t0 + r]
                               : static_cast<const unsigned long long*>(cols.in[a])[t0 + r];
      }
      __syncthreads();
      if (cols.bytes[a] == 4) {
        unsigned* o = static_cast<unsigned*>(cols.out[a]);
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
          const unsigned i = unsigned(k) * PS_NT + tid;
          if (i < tn) o[dst[k]] = unsigned(stage[i]);
        }
      } else {
        unsigned long long* o = static_cast<unsigned long long*>(cols.out[a]);
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
          const unsigned i = unsigned(k) * PS_NT + tid;
          if (i < tn) o[dst[k]] = stage[i];
        }
      }
    }
    __syncthreads();
    // advance the running offsets by this tile's counts; clear for the next
    for (int q = tid; q < P; q += PS_NT) {
      gbase[q] += pc[q];
      pc[q] = 0;
    }
  }
}

// ---- LDS aggregation table shared by passes 3 and 4 -------------------------
// S slots + one extra slot for the key whose bits equal the empty marker.
struct LdsTable {
  unsigned long long* key;
  unsigned long long* cnt;  // 64-bit counts
  unsigned long long* acc;  // [na][S + 1]
  int S;
};

__device__ __forceinline__ LdsTable lds_table(unsigned char* raw, int S, int na) {
  LdsTable t;
  t.key = reinterpret_cast<unsigned long long*>(raw);
  t.cnt = t.key + (S + 1);
  t.acc = t.cnt + (S + 1);
  t.S = S;
  (void)na;
  return t;
}

__device__ __forceinline__ void lds_table_init(const LdsTable& t, const AccSet& accs) {
  for (int s = threadIdx.x; s <= t.S; s += blockDim.x) {
    t.key[s] = HA_EMPTY;
    t.cnt[s] = 0;
    for (int a = 0; a < accs.n; a++)
      t.acc[size_t(a) * (t.S + 1) + s] = acc_identity(accs.a[a].kind, accs.a[a].vtype);
  }
}

// slot of key bits kb (inserting it), or -1 when the table is full
__device__ __forceinline__ int lds_slot(const LdsTable& t, unsigned long long kb,
                                        unsigned long long h) {
  if (kb == HA_EMPTY) return t.S;
  int s = int(h & (unsigned long long)(t.S - 1));
  for (int p = 0; p < t.S; p++) {
    const unsigned long long cur = t.key[s];
    if (cur == kb) return s;
    if (cur == HA_EMPTY) {
      const unsigned long long old = atomicCAS(&t.key[s], HA_EMPTY, kb);
      if (old == HA_EMPTY || old == kb) return s;
    }
    s = (s + 1) & (t.S - 1);
  }
  return -1;
}

// Writes the occupied slots as entries {key, count, acc...}; returns count.
__device__ __forceinline__ void lds_table_flush(const LdsTable& t, int na, int stride,
                                                unsigned long long* out, unsigned* out_n,
                                                unsigned* lds_counter) {
  for (int s = threadIdx.x; s <= t.S; s += blockDim.x) {
    const bool used = (s == t.S) ? t.cnt[s] != 0 : t.key[s] != HA_EMPTY;
    if (!used) continue;
    const unsigned j = atomicAdd(lds_counter, 1u);
    unsigned long long* e = out + size_t(j) * stride;
    e[0] = s == t.S ? HA_EMPTY : t.key[s];
    e[1] = t.cnt[s];
    for (int a = 0; a < na; a++) e[2 + a] = t.acc[size_t(a) * (t.S + 1) + s];
  }
  __syncthreads();
  if (threadIdx.x == 0) *out_n = *lds_counter;
}

// Combines two accumulator words (the same operation as acc_merge, on
// registers); NaN is never selected by max/min.
__device__ __forceinline__ unsigned long long acc_combine(unsigned long long w,
                                                          unsigned long long x, int kind,
                                                          int vtype) {
  switch (vtype) {
    case FHIP_F64: {
      double a, b;
      __builtin_memcpy(&a, &w, 8);
      __builtin_memcpy(&b, &x, 8);
      double r = kind == K_SUM ? a + b
               : kind == K_MAX ? (a < b || a != a ? b : a)
                               : (a > b || a != a ? b : a);
      if (kind != K_SUM && b != b) r = a;  // NaN is never selected
      __builtin_memcpy(&w, &r, 8);
      return w;
    }
    case FHIP_F32: {
      float a, b;
      unsigned ua = unsigned(w), ub = unsigned(x);
      __builtin_memcpy(&a, &ua, 4);
      __builtin_memcpy(&b, &ub, 4);
      float r = kind == K_SUM ? a + b
              : kind == K_MAX ? (a < b || a != a ? b : a)
                              : (a > b || a != a ? b : a);
      if (kind != K_SUM && b != b) r = a;
      unsigned ur;
      __builtin_memcpy(&ur, &r, 4);
      return ur;
    }
    case FHIP_I32: {
      int a = int(unsigned(w)), b = int(unsigned(x));
      int r = kind == K_SUM ? int(unsigned(a) + unsigned(b)) : kind == K_MAX ? (a < b ? b : a) : (a > b ? b : a);
      return unsigned(r);
    }
    case FHIP_U32: {
      unsigned a = unsigned(w), b = unsigned(x);
      return kind == K_SUM ? a + b : kind == K_MAX ? (a < b ? b : a) : (a > b ? b : a);
    }
    case FHIP_I64: {
      long long a = (long long)w, b = (long long)x;
      long long r = kind == K_SUM ? (long long)(w + x) : kind == K_MAX ? (a < b ? b : a) : (a > b ? b : a);
      return (unsigned long long)r;
    }
    default:
      return kind == K_SUM ? w + x : kind == K_MAX ? (w < x ? x : w) : (w > x ? x : w);
  }
}

// Reduces an accumulator word over the lanes in `act` (all lanes end with
// the result); used when a whole wave updates the same LDS slot.
__device__ __forceinline__ unsigned long long wave_reduce_acc(unsigned long long w, int kind,
                                                              int vtype, unsigned long long act,
                                                              bool ok) {
  if (!ok) w = acc_identity(kind, vtype);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) w = acc_combine(w, __shfl_xor(w, o, 64), kind, vtype);
  (void)act;
  return w;
}

struct Item {
  unsigned long long beg, end;
};

// ---- pass 3: aggregate one chunk of one partition in LDS --------------------
// NA = accs.n, 0 to 4: the row registers are sized to the accumulators
// loaded; up to two accumulators two 1024-thread workgroups share a CU.
template <class K, int NA>
__global__ __launch_bounds__(PA_NT, NA <= 2 ? 8 : 4) void pa_agg_kernel(
    const Item* __restrict__ items, const K* __restrict__ keys, AccSet accs, int S, int stride,
    unsigned long long* __restrict__ partial, unsigned* __restrict__ partial_n,
    int* __restrict__ overflow) {
  extern __shared__ unsigned long long lds_u64[];
  unsigned char* lds_raw = reinterpret_cast<unsigned char*>(lds_u64);
  __shared__ unsigned counter;
  __shared__ int full;
  constexpr int na = NA;
  constexpr int NA1 = NA > 0 ? NA : 1;
  constexpr int AGG_UNROLL = NA <= 2 ? 2 : 1;
  LdsTable t = lds_table(lds_raw, S, na);
  lds_table_init(t, accs);
  if (threadIdx.x == 0) {
    counter = 0;
    full = 0;
  }
  __syncthreads();
  const Item it = items[blockIdx.x];
  // the loop bound is tested on the wave's first row so that every lane of a
  // wave stays active through the ballots and shuffles below
  const size_t wave_first = it.beg + (threadIdx.x & ~63u);
  for (size_t w0 = wave_first; w0 < it.end; w0 += size_t(PA_NT) * PA_UNROLL) {
    const size_t r0 = w0 + __lane_id();
    unsigned long long kbv[PA_UNROLL];
    unsigned long long vv[PA_UNROLL][NA1];
#pragma unroll
    for (int u = 0; u < PA_UNROLL; u++) {
      const size_t r = r0 + size_t(u) * PA_NT;
      kbv[u] = r < it.end ? key_bits(keys[r]) : 0ull;
#pragma unroll
      for (int a = 0; a < NA; a++)
        vv[u][a] = r < it.end ? load_bits(accs.a[a].col, accs.a[a].vtype, r) : 0ull;
    }
    const unsigned lane = __lane_id();
#pragma unroll
    for (int u = 0; u < PA_UNROLL; u++) {
      const size_t r = r0 + size_t(u) * PA_NT;
      const bool in = r < it.end;
      const int s = in ? lds_slot(t, kbv[u], mix64(kbv[u])) : -1;
      if (in && s < 0) full = 1;
      const bool ok = s >= 0;
      const unsigned long long act = __ballot(ok);
      if (!act) continue;
      const int leader = __builtin_ffsll((long long)act) - 1;
      const int s0 = __shfl(s, leader, 64);
      if (__ballot(ok && s == s0) == act) {
        // the whole wave updates one slot (a heavy key): reduce, then one
        // atomic per accumulator
        if (lane == unsigned(leader)) atomicAdd(&t.cnt[s0], (unsigned long long)__popcll(act));
#pragma unroll AGG_UNROLL
        for (int a = 0; a < na; a++) {
          const unsigned long long w =
              wave_reduce_acc(vv[u][a], accs.a[a].kind, accs.a[a].vtype, act, ok);
          if (lane == unsigned(leader))
            acc_merge(&t.acc[size_t(a) * (S + 1) + s0], accs.a[a].kind, accs.a[a].vtype, w);
        }
      } else if (ok) {
        atomicAdd(&t.cnt[s], 1ull);
#pragma unroll AGG_UNROLL
        for (int a = 0; a < na; a++)
          acc_merge(&t.acc[size_t(a) * (S + 1) + s], accs.a[a].kind, accs.a[a].vtype, vv[u][a]);
      }
    }
  }
  __syncthreads();
  if (full) {  // the partition has more groups than the table: rebuilt with more partitions
    if (threadIdx.x == 0) {
      partial_n[blockIdx.x] = 0;
      atomicExch(overflow, 1);
    }
    return;
  }
  lds_table_flush(t, na, stride, partial + size_t(blockIdx.x) * (S + 1) * stride,
                  partial_n + blockIdx.x, &counter);
}

// ---- pass 4: merge the chunk partials of one partition ----------------------
__global__ __launch_bounds__(PA_NT) void pa_merge_kernel(
    const unsigned* __restrict__ part_items, const unsigned long long* __restrict__ partial,
    const unsigned* __restrict__ partial_n, AccSet accs, int S, int stride,
    unsigned long long* __restrict__ final_out, unsigned* __restrict__ final_n,
    int* __restrict__ overflow) {
  extern __shared__ unsigned long long lds_u64[];
  unsigned char* lds_raw = reinterpret_cast<unsigned char*>(lds_u64);
  __shared__ unsigned counter;
  __shared__ int full;
  const int na = accs.n;
  const unsigned p = blockIdx.x;
  const unsigned i0 = part_items[p], i1 = part_items[p + 1];
  unsigned long long* out = final_out + size_t(p) * (S + 1) * stride;
  const unsigned cap = unsigned(S + 1);
  if (i1 - i0 == 1) {  // one chunk: its partial is final
    const unsigned m = partial_n[i0] < cap ? partial_n[i0] : cap;
    const unsigned long long* src = partial + size_t(i0) * (S + 1) * stride;
    for (unsigned j = threadIdx.x; j < m * unsigned(stride); j += PA_NT) out[j] = src[j];
    if (threadIdx.x == 0) final_n[p] = m;
    return;
  }
  LdsTable t = lds_table(lds_raw, S, na);
  lds_table_init(t, accs);
  if (threadIdx.x == 0) {
    counter = 0;
    full = 0;
  }
  __syncthreads();
  for (unsigned it = i0; it < i1; it++) {
    const unsigned m = partial_n[it] < cap ? partial_n[it] : cap;
    const unsigned long long* src = partial + size_t(it) * (S + 1) * stride;
    for (unsigned j = threadIdx.x; j < m; j += PA_NT) {
      const unsigned long long* e = src + size_t(j) * stride;
      const unsigned long long kb = e[0];
      const int s = lds_slot(t, kb, mix64(kb));
      if (s < 0) {
        full = 1;
        continue;
      }
      atomicAdd(&t.cnt[s], e[1]);
      for (int a = 0; a < na; a++)
        acc_merge(&t.acc[size_t(a) * (S + 1) + s], accs.a[a].kind, accs.a[a].vtype, e[2 + a]);
    }
  }
  __syncthreads();
  if (full) {
    if (threadIdx.x == 0) {
      final_n[p] = 0;
      atomicExch(overflow, 1);
    }
    return;
  }
  lds_table_flush(t, na, stride, out, final_n + p, &counter);
}

// ---- hot keys: heavy keys of skewed data, aggregated without partitioning --
// The heavy keys come from a sample (at most 80 % of the Sh slots); hot_keys
// holds them at their two-choice slots (HA_EMPTY elsewhere).  Every workgroup walks a
// contiguous chunk of HOT_CHUNK rows: a row with a heavy key is aggregated
// into the workgroup's LDS table, any other row is appended to the
// workgroup's region of the cold buffers (rows r0 .. r0 + seg_n - 1 for a
// chunk starting at row r0; each wave reserves its cold rows of a step with
// one LDS atomic, so the loop has no barrier), and the regions then go
// through the partitioned path as its spans.  The per-workgroup tables share
// the slot layout, so the merge is a column-wise reduction.
constexpr int HOT_NT = 1024;
// rows per thread and step: with no barrier in the loop, one row (1.42 ms
// for 2^28 rows) beat two (1.48), three and four (1.47, 1.48)
constexpr int HOT_UNROLL = 1;
constexpr size_t HOT_CHUNK = size_t(1) << 18;
constexpr int HOT_MAX_LDS = 159 * 1024;  // dynamic LDS of hot_agg_kernel

// The heavy-key table is two-choice (cuckoo) hashed with buckets of two
// slots: a key sits in one of the slots 2*b1, 2*b1+1, 2*b2, 2*b2+1, the
// buckets b1, b2 coming from two multiply-shift hashes (`shift` = 64 -
// log2(Sh / 2)).  A lookup reads those four slots (two 16-byte LDS reads)
// and compares, with no probe loop: every lane of a wave does the same work,
// hit or miss.  A key that the host cannot place stays off the table and
// its rows take the partitioned path.
__host__ __device__ __forceinline__ int hot_bucket1(unsigned long long kb, int shift) {
  return int((kb * 0x9E3779B97F4A7C15ull) >> shift);
}
__host__ __device__ __forceinline__ int hot_bucket2(unsigned long long kb, int shift) {
  return int((kb * 0xC2B2AE3D27D4EB4Full) >> shift);
}

__device__ __forceinline__ int hot_slot(const unsigned long long* hk, int shift,
                                        unsigned long long kb) {
  if (kb == HA_EMPTY) return -1;
  const int s1 = 2 * hot_bucket1(kb, shift), s2 = 2 * hot_bucket2(kb, shift);
  const unsigned long long a0 = hk[s1], a1 = hk[s1 + 1], b0 = hk[s2], b1 = hk[s2 + 1];
  return a0 == kb ? s1 : a1 == kb ? s1 + 1 : b0 == kb ? s2 : b1 == kb ? s2 + 1 : -1;
}

// NA = accs.n, 0 to 4 (a template parameter, so that the row registers are
// sized to the accumulators actually loaded)
// Up to two accumulators, the registers fit two 1024-thread workgroups per
// CU (8 waves per SIMD); the heavy-key table of 2048 slots leaves LDS for both.
// CK: the cold keys as written, K or (for the range path) the low 32 bits.
template <class K, int NA, class CK>
__global__ __launch_bounds__(HOT_NT, NA <= 2 ? 8 : 4) void hot_agg_kernel(
    const K* __restrict__ keys, size_t n, AccSet accs, const unsigned long long* __restrict__ hot_keys,
    int Sh, unsigned long long* __restrict__ partial, CK* __restrict__ cold_keys, Cols cold,
    unsigned long long* __restrict__ cold_n, unsigned long long* __restrict__ seg_n, int cpbits,
    unsigned long long* __restrict__ ccnt, unsigned long long* __restrict__ krange) {
  extern __shared__ unsigned long long lds_u64[];
  __shared__ unsigned cold_cnt;  // rows in this workgroup's cold region
  constexpr int na = NA;
  // the accumulator loops are unrolled for up to two accumulators; wider
  // sets keep the loop (unrolled, they hold too many values in registers)
  constexpr int AGG_UNROLL = NA <= 2 ? 2 : 1;
  // rows per thread and step
  constexpr int U = HOT_UNROLL;
  constexpr int STEP = HOT_NT * U;
  const int shift = 64 - (30 - __builtin_clz(unsigned(Sh)));  // 64 - log2(Sh / 2)
  unsigned long long* hk = lds_u64;        // [Sh]
  unsigned long long* hc = hk + Sh;        // [Sh]
  unsigned long long* ha = hc + Sh;        // [na][Sh]
  // with ccnt: the cold rows' partition counts (2^cpbits partitions) of this
  // workgroup's region, for the partitioned path (it then skips its count pass)
  unsigned* chist = reinterpret_cast<unsigned*>(ha + size_t(na) * Sh);  // [2^cpbits]
  const int CP = ccnt ? 1 << cpbits : 0;
  for (int i = threadIdx.x; i < Sh; i += HOT_NT) {
    hk[i] = hot_keys[i];
    hc[i] = 0;
    for (int a = 0; a < na; a++) ha[size_t(a) * Sh + i] = acc_identity(accs.a[a].kind, accs.a[a].vtype);
  }
  for (int q = threadIdx.x; q < CP; q += HOT_NT) chist[q] = 0;
  if (threadIdx.x == 0) cold_cnt = 0;
  __syncthreads();
  const unsigned tid = threadIdx.x, lane = __lane_id();
  const size_t r0 = size_t(blockIdx.x) * HOT_CHUNK;
  const size_t r1 = r0 + HOT_CHUNK < n ? r0 + HOT_CHUNK : n;
  const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  constexpr int NA1 = NA > 0 ? NA : 1;
  // the rows of the step starting at t (row r >= r1 reads nothing)
  auto load_step = [&](size_t t, K (&kk)[U], unsigned long long (&vals)[U][NA1]) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      const size_t r = t + size_t(u) * HOT_NT + tid;
      kk[u] = r < r1 ? keys[r] : K(0);
#pragma unroll
      for (int a = 0; a < NA; a++)
        vals[u][a] = r < r1 ? load_bits(accs.a[a].col, accs.a[a].vtype, r) : 0ull;
    }
  };
  // the smallest and largest cold key (order-preserving bits: the sign bit
  // flipped for signed keys), for krange; the heavy keys' range is known
  constexpr unsigned long long KFLIP = std::is_signed<K>::value ? (1ull << 63) : 0ull;
  unsigned long long kmn = ~0ull, kmx = 0ull;
  for (size_t t0 = r0; t0 < r1; t0 += STEP) {
    K kv[U];
    unsigned long long vv[U][NA1];
    load_step(t0, kv, vv);
    int sl[U];
    unsigned long long cm[U];  // lanes of the wave with a cold row, per row of the step
    unsigned ncold = 0;
#pragma unroll
    for (int u = 0; u < U; u++) {
      const size_t r = t0 + size_t(u) * HOT_NT + tid;
      const bool in = r < r1;
      sl[u] = in ? hot_slot(hk, shift, key_bits(kv[u])) : -1;
      cm[u] = __ballot(in && sl[u] < 0);
      ncold += unsigned(__popcll(cm[u]));
      const bool ok = sl[u] >= 0;
      const unsigned long long act = __ballot(ok);
      if (!act) continue;
      const int leader = __builtin_ffsll((long long)act) - 1;
      const int s0 = __shfl(sl[u], leader, 64);
      if (__ballot(ok && sl[u] == s0) == act) {
        if (lane == unsigned(leader)) atomicAdd(&hc[s0], (unsigned long long)__popcll(act));
#pragma unroll AGG_UNROLL
        for (int a = 0; a < na; a++) {
          const unsigned long long w = wave_reduce_acc(vv[u][a], accs.a[a].kind, accs.a[a].vtype, act, ok);
          if (lane == unsigned(leader)) acc_merge(&ha[size_t(a) * Sh + s0], accs.a[a].kind, accs.a[a].vtype, w);
        }
      } else if (ok) {
        atomicAdd(&hc[sl[u]], 1ull);
#pragma unroll AGG_UNROLL
        for (int a = 0; a < na; a++)
          acc_merge(&ha[size_t(a) * Sh + sl[u]], accs.a[a].kind, accs.a[a].vtype, vv[u][a]);
      }
    }
    // cold rows of this step: one LDS reservation per wave in the
    // workgroup's region (no barrier, so the waves drift apart freely)
    if (ncold == 0) continue;
    unsigned base = 0;
    if (lane == 0) base = atomicAdd(&cold_cnt, ncold);
    base = unsigned(__shfl(int(base), 0, 64));
#pragma unroll
    for (int u = 0; u < U; u++) {
      if ((cm[u] >> lane) & 1ull) {
        const unsigned long long ob = key_bits(kv[u]) ^ KFLIP;
        kmn = ob < kmn ? ob : kmn;
        kmx = ob > kmx ? ob : kmx;
        const size_t pos = r0 + base + unsigned(__popcll(cm[u] & lt));
        if (CP) atomicAdd(&chist[part_of(mix64(key_bits(kv[u])), cpbits)], 1u);
        cold_keys[pos] = CK(kv[u]);
#pragma unroll
        for (int a = 0; a < NA; a++) {
          if (cold.bytes[a] == 4) static_cast<unsigned*>(cold.out[a])[pos] = unsigned(vv[u][a]);
          else static_cast<unsigned long long*>(cold.out[a])[pos] = vv[u][a];
        }
      }
      base += unsigned(__popcll(cm[u]));
    }
  }
  // the workgroup's cold-key range goes to krange[2b], krange[2b + 1] (no
  // atomics: the host reduces it with the regions' row counts)
  __shared__ unsigned long long s_rng[2 * (HOT_NT / 64)];
  if (krange) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long a = __shfl_xor(kmn, o, 64), b = __shfl_xor(kmx, o, 64);
      kmn = a < kmn ? a : kmn;
      kmx = b > kmx ? b : kmx;
    }
    if (lane == 0) {
      s_rng[2 * (tid >> 6)] = kmn;
      s_rng[2 * (tid >> 6) + 1] = kmx;
    }
  }
  __syncthreads();
  if (tid == 0) {
    if (krange) {
      for (int w = 1; w < HOT_NT / 64; w++) {
        kmn = s_rng[2 * w] < kmn ? s_rng[2 * w] : kmn;
        kmx = s_rng[2 * w + 1] > kmx ? s_rng[2 * w + 1] : kmx;
      }
      krange[2 * blockIdx.x] = kmn;
      krange[2 * blockIdx.x + 1] = kmx;
    }
    seg_n[blockIdx.x] = cold_cnt;
    if (cold_cnt) atomicAdd(cold_n, (unsigned long long)cold_cnt);
  }
  // partition-major, like the count kernel's matrix (one span per workgroup)
  for (int q = tid; q < CP; q += HOT_NT) ccnt[size_t(q) * gridDim.x + blockIdx.x] = chist[q];
  unsigned long long* out = partial + size_t(blockIdx.x) * size_t(1 + na) * Sh;
  for (int i = threadIdx.x; i < Sh; i += HOT_NT) {
    out[i] = hc[i];
    for (int a = 0; a < na; a++) out[size_t(1 + a) * Sh + i] = ha[size_t(a) * Sh + i];
  }
}

// First level of the hot-table reduction: workgroup (x, y) folds the slots
// 64x .. 64x+63 of the per-workgroup tables y*per .. y*per+per-1 into table
// y of `out` (same layout), so the final merge reads gridDim.y tables.
__global__ __launch_bounds__(256) void hot_fold_kernel(
    const unsigned long long* __restrict__ partial, int nparts, int per, int Sh, AccSet accs,
    unsigned long long* __restrict__ out) {
  __shared__ unsigned long long red[4][64][1 + HA_MAX_ACC];
  const int na = accs.n;
  const int slot = blockIdx.x * 64 + int(threadIdx.x & 63);
  const int part = int(threadIdx.x >> 6);
  const int g0 = int(blockIdx.y) * per, g1 = g0 + per < nparts ? g0 + per : nparts;
  unsigned long long c = 0, acc[HA_MAX_ACC];
#pragma unroll
  for (int a = 0; a < HA_MAX_ACC; a++) acc[a] = a < na ? acc_identity(accs.a[a].kind, accs.a[a].vtype) : 0ull;
  if (slot < Sh) {
    for (int g = g0 + part; g < g1; g += 4) {
      const unsigned long long* t = partial + size_t(g) * size_t(1 + na) * Sh;
      c += t[slot];
      for (int a = 0; a < na; a++)
        acc[a] = acc_combine(acc[a], t[size_t(1 + a) * Sh + slot], accs.a[a].kind, accs.a[a].vtype);
    }
  }
  red[part][threadIdx.x & 63][0] = c;
  for (int a = 0; a < na; a++) red[part][threadIdx.x & 63][1 + a] = acc[a];
  __syncthreads();
  if (part != 0 || slot >= Sh) return;
  for (int q = 1; q < 4; q++) {
    c += red[q][threadIdx.x][0];
    for (int a = 0; a < na; a++)
      acc[a] = acc_combine(acc[a], red[q][threadIdx.x][1 + a], accs.a[a].kind, accs.a[a].vtype);
  }
  unsigned long long* t = out + size_t(blockIdx.y) * size_t(1 + na) * Sh;
  t[slot] = c;
  for (int a = 0; a < na; a++) t[size_t(1 + a) * Sh + slot] = acc[a];
}

// Column-wise reduction of the per-workgroup hot tables (in workgroup order);
// every used slot becomes one entry {key, count, acc...} of `out`.
__global__ __launch_bounds__(256) void hot_merge_kernel(
    const unsigned long long* __restrict__ partial, int nparts, int Sh, AccSet accs,
    const unsigned long long* __restrict__ hot_keys, int stride,
    unsigned long long* __restrict__ out, unsigned* __restrict__ out_n) {
  __shared__ unsigned long long red[4][64][1 + HA_MAX_ACC];
  const int na = accs.n;
  const int slot = blockIdx.x * 64 + int(threadIdx.x & 63);
  const int part = int(threadIdx.x >> 6);
  unsigned long long c = 0, acc[HA_MAX_ACC];
#pragma unroll
  for (int a = 0; a < HA_MAX_ACC; a++) acc[a] = a < na ? acc_identity(accs.a[a].kind, accs.a[a].vtype) : 0ull;
  if (slot < Sh) {
    for (int g = part; g < nparts; g += 4) {
      const unsigned long long* t = partial + size_t(g) * size_t(1 + na) * Sh;
      c += t[slot];
      for (int a = 0; a < na; a++)
        acc[a] = acc_combine(acc[a], t[size_t(1 + a) * Sh + slot], accs.a[a].kind, accs.a[a].vtype);
    }
  }
  red[part][threadIdx.x & 63][0] = c;
  for (int a = 0; a < na; a++) red[part][threadIdx.x & 63][1 + a] = acc[a];
  __syncthreads();
  if (part != 0 || slot >= Sh) return;
  for (int q = 1; q < 4; q++) {
    c += red[q][threadIdx.x][0];
    for (int a = 0; a < na; a++)
      acc[a] = acc_combine(acc[a], red[q][threadIdx.x][1 + a], accs.a[a].kind, accs.a[a].vtype);
  }
  if (c == 0) return;
  const unsigned j = atomicAdd(out_n, 1u);
  unsigned long long* e = out + size_t(j) * stride;
  e[0] = hot_keys[slot];
  e[1] = c;
  for (int a = 0; a < na; a++) e[2 + a] = acc[a];
}

// Range of the group keys of the final tables in the radix sort's order (the
// sign bit of signed keys flipped): range[0] min, range[1] max, updated with
// one atomic each per partition.  The group-key sort then needs no
// histogram readback.
__global__ __launch_bounds__(256) void key_range_kernel(const unsigned long long* __restrict__ w,
                                                        int stride, int S,
                                                        const unsigned* __restrict__ fn, int kb,
                                                        int is_signed,
                                                        unsigned long long* __restrict__ range) {
  __shared__ unsigned long long smin[256], smax[256];
  const int p = blockIdx.x;
  const unsigned m = fn[p];
  const unsigned long long mask = kb == 8 ? ~0ull : (1ull << (8 * kb)) - 1;
  const unsigned long long top = 1ull << (8 * kb - 1);
  unsigned long long lo = ~0ull, hi = 0;
  for (unsigned j = threadIdx.x; j < m; j += 256) {
    unsigned long long b = w[(size_t(p) * (S + 1) + j) * size_t(stride)] & mask;
    if (is_signed) b ^= top;
    lo = b < lo ? b : lo;
    hi = b > hi ? b : hi;
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (unsigned h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      smin[threadIdx.x] = smin[threadIdx.x + h] < smin[threadIdx.x] ? smin[threadIdx.x + h] : smin[threadIdx.x];
      smax[threadIdx.x] = smax[threadIdx.x + h] > smax[threadIdx.x] ? smax[threadIdx.x + h] : smax[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && m) {
    atomicMin(&range[0], smin[0]);
    atomicMax(&range[1], smax[0]);
  }
}

__global__ void final_index_kernel(const unsigned* __restrict__ final_n,
                                   const unsigned long long* __restrict__ final_off, int P, int S,
                                   size_t* __restrict__ ent) {
  const int p = blockIdx.x;
  const unsigned m = final_n[p];
  const unsigned long long o = final_off[p];
  for (unsigned j = threadIdx.x; j < m; j += blockDim.x)
    ent[o + j] = size_t(p) * (S + 1) + j;
}

// g_dev (optional): a group count on the device; at most that many of the
// G groups are written
template <class K>
__global__ void gather_group_keys(const unsigned long long* __restrict__ w, int stride,
                                  const size_t* __restrict__ ent, size_t G, K* __restrict__ keys,
                                  const unsigned* __restrict__ g_dev = nullptr) {
  const size_t st = size_t(gridDim.x) * blockDim.x;
  if (g_dev) G = G < size_t(*g_dev) ? G : size_t(*g_dev);
  for (size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x; g < G; g += st)
    keys[g] = K(w[ent[g] * stride]);
}

// The aggregations of one fused group-by, for the final write.
// Dense key range (ordered key bits in [lo, lo + R)): flag[b - lo] = 1 for
// every group's key; the inclusive scan of the flags then gives each key
// its rank among the groups, which replaces the sort of the group keys.
__global__ void dense_mark_kernel(const unsigned long long* __restrict__ w, int stride,
                                  const size_t* __restrict__ ent, size_t G,
                                  unsigned long long mask, unsigned long long flip,
                                  unsigned long long lo, unsigned* __restrict__ flag) {
  const size_t st = size_t(gridDim.x) * blockDim.x;
  for (size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x; g < G; g += st)
    flag[((w[ent[g] * stride] & mask) ^ flip) - lo] = 1u;
}

// ent_sorted[rank of group g's key] = ent[g] (ranks from the inclusive scan)
__global__ void dense_order_kernel(const unsigned long long* __restrict__ w, int stride,
                                   const size_t* __restrict__ ent, size_t G,
                                   unsigned long long mask, unsigned long long flip,
                                   unsigned long long lo, const unsigned* __restrict__ incl,
                                   size_t* __restrict__ ent_sorted) {
  const size_t st = size_t(gridDim.x) * blockDim.x;
  for (size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x; g < G; g += st) {
    const size_t e = ent[g];
    ent_sorted[incl[((w[e * stride] & mask) ^ flip) - lo] - 1u] = e;
  }
}

struct EmitSet {
  int n;
  int op[16];
  int vtype[16];
  int slot[16];  // accumulator of the aggregation (count has none)
  void* out[16];
};

// Writes every aggregate in sorted-group order (one thread per group reads
// the group's entry once for all of them).
__global__ void emit_kernel(const unsigned long long* __restrict__ w, int stride,
                            const size_t* __restrict__ ent, size_t G, EmitSet es,
                            const unsigned* __restrict__ g_dev = nullptr) {
  const size_t st = size_t(gridDim.x) * blockDim.x;
  if (g_dev) G = G < size_t(*g_dev) ? G : size_t(*g_dev);
  for (size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x; g < G; g += st) {
    const unsigned long long* e = w + ent[g] * stride;
    const unsigned long long cnt = e[1];
    for (int i = 0; i < es.n; i++) {
      const int op = es.op[i], vtype = es.vtype[i];
      void* out = es.out[i];
      if (op == OP_COUNT) {
        static_cast<unsigned long long*>(out)[g] = cnt;
        continue;
      }
      const unsigned long long a = e[2 + es.slot[i]];
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
}

int grid_for(size_t n, int per_cu = 8) {
  return int(std::max<size_t>(1, std::min<size_t>(ceil_div(n, 256), size_t(fill_blocks(256, per_cu)))));
}

template <class K>
void launch_pa(const void* keys, size_t n, int pbits, size_t spans, int span_tiles, const Cols& cols,
               unsigned long long* cnt, unsigned long long* incl, void* pkeys, hipStream_t st,
               int pass, const unsigned long long* seg_n, size_t seg_stride) {
  const int P = 1 << pbits;
  if (pass == 1) {
    pa_count_kernel<K><<<unsigned(spans), PS_NT, P * 4, st>>>((const K*)keys, n, pbits, spans,
                                                               span_tiles, cnt, seg_n, seg_stride);
  } else {
    if (P <= PS_WIDE_MAX_P) {
      const size_t tile = size_t(PS_NT) * PS_ITEMS_WIDE;
      const size_t lds = tile * 8 + size_t(P) * 16 + tile * 2;
      pa_scatter_kernel<K, PS_ITEMS_WIDE><<<unsigned(spans), PS_NT, lds, st>>>(
          (const K*)keys, n, pbits, spans, span_tiles, cols, cnt, incl, (K*)pkeys, seg_n, seg_stride);
    } else {
      const size_t lds = size_t(PS_TILE) * 8 + size_t(P) * 16 + size_t(PS_TILE) * 2;
      pa_scatter_kernel<K, PS_ITEMS><<<unsigned(spans), PS_NT, lds, st>>>(
          (const K*)keys, n, pbits, spans, span_tiles, cols, cnt, incl, (K*)pkeys, seg_n, seg_stride);
    }
  }
  FHIP_CHECK(hipGetLastError());
}

}  // namespace

void group_by(int ktype, const void* col, size_t* idx, size_t n, size_t* split, size_t* nsplit,
              void* keys_out, hipStream_t st);
void prefix_sum(int type, const void* in, void* out, size_t n, hipStream_t st);
void grouped_agg(int op, int vtype, const void* col, size_t col_n, const size_t* gidx, size_t n,
                 const size_t* split, size_t nsplit, const size_t* nulls, size_t nn, void* out,
                 hipStream_t st);

// Heavy-key candidates.  The sample is m rows in m / HOT_GROUP evenly spaced
// groups of HOT_GROUP consecutive rows (few memory pages, so the reads do
// not wait on address translation).  A key counts once per group that holds
// it, so a run of equal keys (sorted or clustered data) does not look heavy;
// for keys of frequency p << 1 / HOT_GROUP the count expects m * p, as for m
// single rows.  The counts go to an open-addressing table of 2m slots
// (linear probing from a mixed hash; at most half the slots fill, so every
// probe ends), then the keys counted at least HOT_MIN_SEEN times are listed
// as (count, key) pairs after a leading entry count.  The list is in
// arrival order; the host sorts it by (count descending, key ascending),
// which fixes the plan.  Every slot also counts the sample rows that hold
// its key (not once per group), so the plan measures how many of the
// sample's rows the kept keys cover in rows, as its threshold is stated.
constexpr unsigned HOT_MIN_SEEN = 4;
constexpr unsigned HOT_GROUP = 64;  // the wave width (or a divisor of it)

__global__ void hot_count_init_kernel(unsigned long long* __restrict__ tk, unsigned* __restrict__ tc,
                                      unsigned* __restrict__ tr, unsigned size,
                                      unsigned long long* __restrict__ list) {
  const unsigned stride = gridDim.x * blockDim.x;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < size; i += stride) {
    tk[i] = HA_EMPTY;
    tc[i] = 0;
    tr[i] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) list[0] = 0;
}

// gstep: rows between the starts of consecutive sample groups (>= HOT_GROUP)
template <class K>
__global__ void hot_count_kernel(const K* __restrict__ keys, size_t gstep, size_t m,
                                 unsigned long long* __restrict__ tk, unsigned* __restrict__ tc,
                                 unsigned* __restrict__ tr, unsigned mask) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  const unsigned lane = __lane_id();
  const unsigned o = lane % HOT_GROUP;  // position in the group (groups never straddle waves)
  // the loop runs a whole number of times for every lane of a wave, so the
  // ballots and shuffles below see all of them
  const size_t m_up = (m + stride - 1) / stride * stride;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m_up; i += stride) {
    const size_t row = (i / HOT_GROUP) * gstep + o;
    const unsigned long long kb = i < m ? key_bits(keys[row]) : HA_EMPTY;
    // only the group's first lane with a key counts it
    bool first = true;
    const int lo = int(unsigned(kb)), hi = int(unsigned(kb >> 32));
    for (unsigned d = 1; d < HOT_GROUP; d++) {
      const int src = int(lane) - int(d);
      const unsigned olo = unsigned(__shfl(lo, src < 0 ? 0 : src, 64));
      const unsigned ohi = unsigned(__shfl(hi, src < 0 ? 0 : src, 64));
      if (o >= d && ((unsigned long long)ohi << 32 | olo) == kb) first = false;
    }
    // the empty marker is never a heavy key; every lane with a key finds its
    // slot (rows), the group's first one also counts the group (presence)
    const bool has = kb != HA_EMPTY;
    const bool valid = has && first;
    unsigned h = unsigned(mix64(kb)) & mask;
    while (has) {
      // a slot's key never changes once set: most lanes of a skewed sample
      // find theirs with a plain read and leave the CAS to the first one
      const unsigned long long cur = __hip_atomic_load(&tk[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (cur == kb) break;
      if (cur == HA_EMPTY) {
        const unsigned long long prev = atomicCAS(&tk[h], HA_EMPTY, kb);
        if (prev == HA_EMPTY || prev == kb) break;
      }
      h = (h + 1) & mask;
    }
    // lanes with the same slot add their counts with one atomic each
    const unsigned long long firsts = __ballot(valid);
    unsigned long long pending = __ballot(has);
    while (pending) {
      const int leader = __builtin_ffsll((long long)pending) - 1;
      const unsigned hs = unsigned(__shfl(int(h), leader, 64));
      const unsigned long long same = __ballot(has && h == hs) & pending;
      if (lane == unsigned(leader)) {
        const unsigned c = unsigned(__popcll(same & firsts));
        if (c) atomicAdd(&tc[hs], c);
        atomicAdd(&tr[hs], unsigned(__popcll(same)));
      }
      pending &= ~same;
    }
  }
}

// list entries are (presence count, key, sample rows) triples
__global__ void hot_list_kernel(const unsigned long long* __restrict__ tk,
                                const unsigned* __restrict__ tc, const unsigned* __restrict__ tr,
                                unsigned size, unsigned long long* __restrict__ list,
                                unsigned cap) {
  const unsigned stride = gridDim.x * blockDim.x;
  for (unsigned s = blockIdx.x * blockDim.x + threadIdx.x; s < size; s += stride) {
    const unsigned c = tc[s];
    if (c < HOT_MIN_SEEN) continue;
    const unsigned long long j = atomicAdd(&list[0], 1ull);
    if (j < cap) {
      list[1 + 3 * j] = c;
      list[2 + 3 * j] = tk[s];
      list[3 + 3 * j] = tr[s];
    }
  }
}

// The heavy keys of a skewed key column: keys found in at least 4 of the
// 2^11 groups of a 2^17-row sample (expected frequency >= 2^-15), at most
// 80 % of Sh of them (the most frequent first), placed in a two-choice
// table of Sh slots.  Not used when their counts cover less than
// groupby.hot_min_pct percent of the sample.
struct HotPlan {
  bool used = false;
  int Sh = 0;
  std::vector<unsigned long long> table;
};

// Two-choice placement of the candidates `runs` ((count, key bits), the
// heaviest first) into a table of Sh slots: at most 80 % of the slots are
// filled, a key goes to a free slot of either bucket, or displaces a key of
// the bucket it was not just evicted from, which then moves on; a key still
// unplaced after the displacement limit is left out (its rows take the
// partitioned path).
HotPlan place_hot(std::vector<std::pair<size_t, unsigned long long>>& runs, int Sh) {
  HotPlan h;
  if (runs.size() > size_t(Sh) * 4 / 5) runs.resize(size_t(Sh) * 4 / 5);
  h.Sh = Sh;
  h.table.assign(size_t(Sh), HA_EMPTY);
  const int shift = 64 - (__builtin_ctz(unsigned(Sh)) - 1);
  for (auto& r : runs) {
    unsigned long long k = r.second;
    int from = -1;  // the bucket k was evicted from
    for (int kick = 0; kick < 256; kick++) {
      const int c1 = hot_bucket1(k, shift), c2 = hot_bucket2(k, shift);
      int placed = -1;
      for (int s : {2 * c1, 2 * c1 + 1, 2 * c2, 2 * c2 + 1})
        if (h.table[size_t(s)] == HA_EMPTY) {
          placed = s;
          break;
        }
      if (placed >= 0) {
        h.table[size_t(placed)] = k;
        k = HA_EMPTY;
        break;
      }
      const int eb = (from == c1) ? c2 : c1;
      std::swap(k, h.table[size_t(2 * eb + (kick & 1))]);
      from = eb;
    }
  }
  h.used = true;
  return h;
}

// Table slots of the heavy-key pass for a group-by whose LDS tables have S
// slots: 2048 by default (the 80 % the plan fills hold the sampled heavy
// keys of skewed data -- about 1,900 keys of 2^20 Zipf(1.1) values reach 4
// groups of a 2^17-row sample; the 1,638 most counted are kept -- and two
// workgroups' tables fit one CU's LDS); 0 when too small
int hot_slots(int S) {
  long long want = tune_get("groupby.hot", std::min(S, 2048));
  if (want < 256) return 0;
  int Sh = int(std::min<long long>(want, S));
  while (Sh & (Sh - 1)) Sh &= Sh - 1;
  return Sh;
}

// The heavy-key plan from known heavy keys: the same coverage rule as the
// sampled plan (the kept keys must hold groupby.hot_min_pct % of the rows)
// The plan depends on the list, the slot count and the coverage rule only, so
// a list with a cache (a column model's) is placed once per (Sh, rule).
HotPlan plan_hot_known(const HeavyKeys& hk, size_t n, int S) {
  const int Sh = hot_slots(S);
  if (!Sh || hk.n == 0 || hk.total == 0 || n < (size_t(1) << 16)) return HotPlan();
  const long long min_pct = tune_get("groupby.hot_min_pct", 30);
  std::unique_lock<std::mutex> lock;
  if (hk.cache) {
    lock = std::unique_lock<std::mutex>(hk.cache->mu);
    const HotPlanCache& c = *hk.cache;
    if (c.valid && c.Sh == Sh && c.min_pct == min_pct) {
      HotPlan h;
      if (c.used) {
        h.used = true;
        h.Sh = c.Sh;
        h.table = c.table;
      }
      return h;
    }
  }
  std::vector<std::pair<size_t, unsigned long long>> runs(hk.n);
  for (size_t j = 0; j < hk.n; j++) runs[j] = {size_t(hk.counts[j]), hk.keys[j]};
  const size_t keep = std::min(runs.size(), size_t(Sh) * 4 / 5);
  size_t covered = 0;
  for (size_t j = 0; j < keep; j++) covered += runs[j].first;
  HotPlan h;
  if (covered * 100 >= size_t(std::max(0LL, min_pct)) * hk.total) h = place_hot(runs, Sh);
  if (hk.cache) {
    HotPlanCache& c = *hk.cache;
    c.valid = true;
    c.Sh = Sh;
    c.min_pct = min_pct;
    c.used = h.used;
    c.table = h.table;
  }
  return h;
}

HotPlan plan_hot(int ktype, const void* keys, size_t n, int S, hipStream_t st) {
  HotPlan h;
  const int Sh = hot_slots(S);
  // sample of 2^groupby.hot_sample rows (default 2^17)
  const int mbits = int(std::max(12LL, std::min(20LL, tune_get("groupby.hot_sample", 17))));
  const size_t m = size_t(1) << mbits;
  if (!Sh || n < 16 * m) return h;
  // n >= 16 m, so the groups (HOT_GROUP rows every gstep rows) never overlap
  const size_t gstep = n / (m / HOT_GROUP);
  const unsigned tsize = unsigned(2 * m), cap = unsigned(m / HOT_MIN_SEEN);
  DBuf tk(size_t(tsize) * 8), tc(size_t(tsize) * 4), trb(size_t(tsize) * 4),
      lb((1 + 3 * size_t(cap)) * 8);
  unsigned long long* list = lb.as<unsigned long long>();
  unsigned long long* tkp = tk.as<unsigned long long>();
  unsigned* tcp = tc.as<unsigned>();
  unsigned* trp = trb.as<unsigned>();
  hot_count_init_kernel<<<grid_for_n(tsize), 256, 0, st>>>(tkp, tcp, trp, tsize, list);
  switch (ktype) {
    case FHIP_I32: hot_count_kernel<int32_t><<<grid_for_n(m), 256, 0, st>>>((const int32_t*)keys, gstep, m, tkp, tcp, trp, tsize - 1); break;
    case FHIP_U32: hot_count_kernel<uint32_t><<<grid_for_n(m), 256, 0, st>>>((const uint32_t*)keys, gstep, m, tkp, tcp, trp, tsize - 1); break;
    case FHIP_I64: hot_count_kernel<int64_t><<<grid_for_n(m), 256, 0, st>>>((const int64_t*)keys, gstep, m, tkp, tcp, trp, tsize - 1); break;
    default: hot_count_kernel<uint64_t><<<grid_for_n(m), 256, 0, st>>>((const uint64_t*)keys, gstep, m, tkp, tcp, trp, tsize - 1); break;
  }
  hot_list_kernel<<<grid_for_n(tsize), 256, 0, st>>>(tkp, tcp, trp, tsize, list, cap);
  FHIP_CHECK(hipGetLastError());
  // the count and the first 4096 candidates in one copy (skewed data has
  // about a thousand), the rest only when there are more
  const size_t first = std::min<size_t>(cap, 4096);
  // (pinned slot 2: its next user, the group-by's final readback, comes
  // after this copy has been waited for)
  unsigned long long* lst =
      static_cast<unsigned long long*>(pinned_scratch(2, (1 + 3 * size_t(cap)) * 8));
  FHIP_CHECK(hipMemcpyAsync(lst, list, (1 + 3 * first) * 8, hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  const size_t nc = std::min<size_t>(size_t(lst[0]), cap);
  if (nc > first) {
    FHIP_CHECK(hipMemcpyAsync(lst + 1 + 3 * first, list + 1 + 3 * first,
                              3 * (nc - first) * 8, hipMemcpyDeviceToHost, st));
    FHIP_CHECK(hipStreamSynchronize(st));
  }
  // (count, key) by count descending, then key ascending: a counting sort on
  // the count (at most m / HOT_GROUP), then the keys of each count sorted;
  // the sample rows of every key ride along for the coverage below
  std::vector<std::pair<size_t, unsigned long long>> runs(nc);
  std::vector<unsigned long long> rows_of(nc);
  {
    size_t maxc = 0;
    for (size_t j = 0; j < nc; j++) maxc = std::max<size_t>(maxc, size_t(lst[1 + 3 * j]));
    std::vector<size_t> start(maxc + 2, 0);  // start[c]: first position of count c
    for (size_t j = 0; j < nc; j++) start[maxc - size_t(lst[1 + 3 * j]) + 1]++;
    for (size_t c = 1; c < start.size(); c++) start[c] += start[c - 1];
    std::vector<std::pair<unsigned long long, unsigned long long>> ks(nc);  // (key, rows)
    for (size_t j = 0; j < nc; j++)
      ks[start[maxc - size_t(lst[1 + 3 * j])]++] = {lst[2 + 3 * j], lst[3 + 3 * j]};
    // start[maxc - c] now ends the run of count c
    size_t b = 0;
    for (size_t r = 0; r <= maxc; r++) {
      const size_t e = start[r];
      if (e - b > 1) std::sort(ks.begin() + b, ks.begin() + e);
      for (size_t j = b; j < e; j++) {
        runs[j] = {maxc - r, ks[j].first};
        rows_of[j] = ks[j].second;
      }
      b = e;
    }
  }
  // two-choice buckets of two slots place keys up to a load near 90 %; the
  // plan fills at most 80 % of the slots.  Sample rows holding a kept key,
  // against the sample's m rows
  const size_t keep = std::min(runs.size(), size_t(Sh) * 4 / 5);
  size_t covered = 0;
  for (size_t j = 0; j < keep; j++) covered += size_t(rows_of[j]);
  const long long min_pct = tune_get("groupby.hot_min_pct", 30);
  if (runs.empty() || covered * 100 < size_t(std::max(0LL, min_pct)) * m) return h;
  return place_hot(runs, Sh);
}

static __global__ void ha_iota_kernel(size_t* __restrict__ p, size_t n) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = i;
}

// More groups than the partitioned tables hold: sort-based group_by plus
// per-column aggregation (same results).
void group_by_agg_sorted(int ktype, const void* keys, size_t n, const fhip_agg* aggs, int naggs,
                         void* keys_out, size_t capacity, size_t* ngroups, hipStream_t st) {
  const int kb = type_bytes(ktype);
  DBuf idx(n * 8), split((n + 1) * 8), gk(n * size_t(kb));
  ha_iota_kernel<<<grid_for_n(n), 256, 0, st>>>(idx.as<size_t>(), n);
  FHIP_CHECK(hipGetLastError());
  size_t ns = 0;
  group_by(ktype, keys, idx.as<size_t>(), n, split.as<size_t>(), &ns, gk.get(), st);
  const size_t G = ns - 1;
  *ngroups = G;
  FHIP_REQUIRE(G <= capacity, FHIP_ERR_TOO_LARGE,
               "group_by_agg: output capacity " + std::to_string(capacity) + " < groups " +
                   std::to_string(G));
  FHIP_CHECK(hipMemcpyAsync(keys_out, gk.get(), G * size_t(kb), hipMemcpyDeviceToDevice, st));
  for (int i = 0; i < naggs; i++)
    grouped_agg(aggs[i].op, aggs[i].op == OP_COUNT ? ktype : aggs[i].vtype,
                aggs[i].op == OP_COUNT ? keys : aggs[i].col, n, idx.as<size_t>(), n,
                split.as<size_t>(), ns, nullptr, 0, aggs[i].out, st);
}

// Sorted-by-key result of a fused group-by; see fhip_dev_group_by_agg.
// heavy (optional): the column's heavy keys known in advance (a model of the
// column); then no sample of the input is counted for the heavy-key table
namespace {

// ---- range path: group keys in a narrow range ------------------------------
// When the heavy-key pass finds all keys within R <= min(rows, 2^21)
// consecutive values, the cold rows are partitioned by key offset instead of
// by hash, and every part's table is an array indexed by the offset:
//   rp_part: each 16384-row tile of a cold region ranks its rows by part
//     (key offset k goes to part (k / 16) mod P, in runs of 16 keys; P a
//     power of two, at most 1024, parts of at most 2^pkb keys) in LDS and
//     writes the 16-bit index within the part and the value columns in part
//     order to the tile's own slots, with the start of every part in the
//     tile (part-major, as the dense join's partition does);
//   rp_agg: workgroup (part, chunk of tiles) aggregates its rows into LDS
//     arrays indexed by the index (no hashing, no key compare, no
//     overflow) and writes them out per chunk;
//   rp_combine: per key, the chunks' arrays are combined (in chunk order)
//     into one entry {key, count, accumulators} and flagged when present;
//     rp_hot then places the heavy keys' entries;
//   the flags' prefix sum ranks the present keys, which are already in key
//   order: no sort of the group keys.
constexpr int RP_NT = 1024;
constexpr int RP_ITEMS = 16;
constexpr int RP_TILE = RP_NT * RP_ITEMS;  // 16384 rows: tile positions fit 16 bits
constexpr int RP_MAXP = 1024;              // parts at most (one counter per thread)
constexpr size_t RP_TPR = HOT_CHUNK / RP_TILE;  // tiles per heavy-key region
static_assert(HOT_CHUNK % RP_TILE == 0, "a heavy-key region is whole tiles");
constexpr int RP_ANT = 512;                // threads of the aggregation kernel
constexpr int RP_AU = 8;                   // rows per lane in flight (aggregation)
constexpr size_t RP_LDS = size_t(80) << 10;  // LDS of one part's arrays (two workgroups a CU)
constexpr unsigned long long RP_MAXR = 1ull << 21;
constexpr int RP_BLK = 4;                  // keys go to the parts in blocks of 2^RP_BLK

// Index of key offset k within its part (ppb = log2 of the part count).
__host__ __device__ __forceinline__ unsigned rp_index(unsigned long long k, int ppb) {
  return unsigned((((k >> RP_BLK) >> ppb) << RP_BLK) | (k & ((1ull << RP_BLK) - 1)));
}

// Tile t holds the `rows` cold rows from row `src` of the cold buffers
// (tiles[t] = src << 14 | rows - 1; every tile lies in one cold region and
// holds at least one row) and writes slots t * RP_TILE .. of the outputs.
template <class K, int NA>
__global__ __launch_bounds__(RP_NT) void rp_part_kernel(
    const K* __restrict__ ckeys, Cols cold, const unsigned long long* __restrict__ tiles,
    size_t ntiles, unsigned long long kmin, int P, unsigned short* __restrict__ lk_out,
    Cols pout, unsigned short* __restrict__ starts) {
  __shared__ __align__(16) unsigned char s_stage[size_t(RP_TILE) * 8];
  __shared__ unsigned s_cnt[RP_MAXP];
  __shared__ unsigned s_tmp[RP_NT / 64];
  __shared__ unsigned s_total;
  // tiles t .. t + per - 1 on one XCD, so neighbouring tiles' start words
  // (which share cache lines) are combined in one L2
  const size_t per = (ntiles + 7) / 8;
  const size_t t = size_t(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (t >= ntiles) return;
  const unsigned long long td = tiles[t];
  const unsigned rows = unsigned(td & 0x3FFFull) + 1;
  const size_t lbase = size_t(td >> 14);  // the tile's first cold row
  const size_t base = t * RP_TILE;        // its output slots
  // loads go through clamped indices (no branch)
  const K* kt = ckeys + lbase;
  for (int p = threadIdx.x; p < RP_MAXP; p += RP_NT) s_cnt[p] = 0;
  // offset k in block k / 16: part (k / 16) mod P, index
  // (k / 16 / P) * 16 + k mod 16 within the part (P a power of two), so
  // every part holds an even share of each stretch of keys (a skewed
  // column's densest stretch would otherwise fill a few parts) and a part's
  // keys still come in runs of 16
  const unsigned long long pmask = (unsigned long long)(P - 1);
  const int ppb = __builtin_ctz(unsigned(P));
  unsigned kl[RP_ITEMS / 2];  // two 16-bit indexes within the part per word
  unsigned dr[RP_ITEMS];      // part << 16 | rank within the part, ~0u past the rows
  {
    unsigned long long kb[RP_ITEMS];
#pragma unroll
    for (int j = 0; j < RP_ITEMS; j++) {
      const unsigned i = unsigned(j) * RP_NT + threadIdx.x;
      // the range spans less than 2^32 keys: the offset is the low 32 bits
      // of the difference (cold keys kept as their low 32 bits give the same)
      kb[j] = (unsigned long long)unsigned(key_bits(kt[i < rows ? i : 0u]) - kmin);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RP_ITEMS; j++) {
      const unsigned i = unsigned(j) * RP_NT + threadIdx.x;
      const unsigned l = rp_index(kb[j], ppb);
      if (j & 1) kl[j >> 1] |= l << 16;
      else kl[j >> 1] = l;
      dr[j] = ~0u;
      if (i < rows) {
        const unsigned d = unsigned((kb[j] >> RP_BLK) & pmask);
        dr[j] = (d << 16) | atomicAdd(&s_cnt[d], 1u);
      }
    }
  }
  __syncthreads();
  {
    const unsigned c = s_cnt[threadIdx.x];
    unsigned all;
    const unsigned e = block_excl_scan<RP_NT>(c, s_tmp, &all);
    s_cnt[threadIdx.x] = e;
    if (int(threadIdx.x) < P) starts[size_t(threadIdx.x) * ntiles + t] = (unsigned short)e;
    if (threadIdx.x == 0) {
      s_total = all;
      starts[size_t(P) * ntiles + t] = (unsigned short)all;
    }
  }
  __syncthreads();
  const unsigned cnt = s_total;
  // the offsets within the parts, then every value column, each staged in
  // part order and written out in one run
  {
    unsigned short* sk = reinterpret_cast<unsigned short*>(s_stage);
#pragma unroll
    for (int j = 0; j < RP_ITEMS; j++)
      if (dr[j] != ~0u)
        sk[s_cnt[dr[j] >> 16] + (dr[j] & 0xFFFFu)] = (unsigned short)(kl[j >> 1] >> ((j & 1) * 16));
    __syncthreads();
    unsigned short* lo = lk_out + base;
    for (unsigned i = threadIdx.x; i < cnt; i += RP_NT) lo[i] = sk[i];
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < NA; a++) {
    if (cold.bytes[a] == 4) {
      const unsigned* src = static_cast<const unsigned*>(cold.in[a]) + lbase;
      unsigned* sv = reinterpret_cast<unsigned*>(s_stage);
      unsigned v[RP_ITEMS];
#pragma unroll
      for (int j = 0; j < RP_ITEMS; j++) {
        const unsigned i = unsigned(j) * RP_NT + threadIdx.x;
        v[j] = src[i < rows ? i : 0u];
      }
#pragma unroll
      for (int j = 0; j < RP_ITEMS; j++)
        if (dr[j] != ~0u) sv[s_cnt[dr[j] >> 16] + (dr[j] & 0xFFFFu)] = v[j];
      __syncthreads();
      unsigned* dst = static_cast<unsigned*>(pout.out[a]) + base;
      for (unsigned i = threadIdx.x; i < cnt; i += RP_NT) dst[i] = sv[i];
    } else {
      const unsigned long long* src = static_cast<const unsigned long long*>(cold.in[a]) + lbase;
      unsigned long long* sv = reinterpret_cast<unsigned long long*>(s_stage);
      unsigned long long v[RP_ITEMS];
#pragma unroll
      for (int j = 0; j < RP_ITEMS; j++) {
        const unsigned i = unsigned(j) * RP_NT + threadIdx.x;
        v[j] = src[i < rows ? i : 0u];
      }
#pragma unroll
      for (int j = 0; j < RP_ITEMS; j++)
        if (dr[j] != ~0u) sv[s_cnt[dr[j] >> 16] + (dr[j] & 0xFFFFu)] = v[j];
      __syncthreads();
      unsigned long long* dst = static_cast<unsigned long long*>(pout.out[a]) + base;
      for (unsigned i = threadIdx.x; i < cnt; i += RP_NT) dst[i] = sv[i];
    }
    __syncthreads();
  }
}

// Workgroup w aggregates part w / C over the tile chunk w % C into LDS
// arrays indexed by the index within the part (count: 32 bits; one 64-bit
// word per accumulator) and writes them to chunk w % C's arrays, part-major
// (pcnt[c * P * 2^pkb + p * 2^pkb + index], pacc likewise per accumulator).
template <int NA>
__global__ __launch_bounds__(RP_ANT) void rp_agg_kernel(
    const unsigned short* __restrict__ lk_in, AccSet accs,
    const unsigned short* __restrict__ starts, size_t ntiles, int P, int C, int pkb,
    unsigned* __restrict__ pcnt, unsigned long long* __restrict__ pacc) {
  extern __shared__ unsigned long long lds_u64[];
  const int PK = 1 << pkb;
  unsigned long long* acc = lds_u64;                          // [NA][PK]
  unsigned* cnt = reinterpret_cast<unsigned*>(acc + size_t(NA) * PK);  // [PK]
  const int p = int(blockIdx.x) / C, c = int(blockIdx.x) % C;
  for (int i = threadIdx.x; i < PK; i += RP_ANT) {
    cnt[i] = 0;
#pragma unroll
    for (int a = 0; a < NA; a++) acc[size_t(a) * PK + i] = acc_identity(accs.a[a].kind, accs.a[a].vtype);
  }
  __syncthreads();
  const size_t chunk = (ntiles + size_t(C) - 1) / size_t(C);
  const size_t t0 = size_t(c) * chunk;
  const size_t t1 = std::min(ntiles, t0 + chunk);
  constexpr int NA1 = NA > 0 ? NA : 1;
  if (t0 < t1)
    segment_rows<RP_TILE, RP_AU>(starts, ntiles, p, t0, t1, threadIdx.x >> 6, RP_ANT / 64,
                                 [&](const size_t* pos, const bool* in) {
                                   unsigned l[RP_AU];
                                   unsigned long long v[NA1][RP_AU];
#pragma unroll
                                   for (int u = 0; u < RP_AU; u++) l[u] = lk_in[pos[u]];
                                   // one branch per column (its width is
                                   // uniform), so that all its loads issue
                                   // together
#pragma unroll
                                   for (int a = 0; a < NA; a++) {
                                     const int vt = accs.a[a].vtype;
                                     if (vt == FHIP_I32 || vt == FHIP_U32 || vt == FHIP_F32) {
                                       const unsigned* c = static_cast<const unsigned*>(accs.a[a].col);
#pragma unroll
                                       for (int u = 0; u < RP_AU; u++) v[a][u] = c[pos[u]];
                                     } else {
                                       const unsigned long long* c =
                                           static_cast<const unsigned long long*>(accs.a[a].col);
#pragma unroll
                                       for (int u = 0; u < RP_AU; u++) v[a][u] = c[pos[u]];
                                     }
                                   }
#pragma unroll
                                   for (int u = 0; u < RP_AU; u++)
                                     if (in[u]) atomicAdd(&cnt[l[u]], 1u);
#pragma unroll
                                   for (int a = 0; a < NA; a++)
                                     acc_merge_rows<RP_AU>(acc + size_t(a) * PK, l, in,
                                                           accs.a[a].kind, accs.a[a].vtype, v[a]);
                                 });
  __syncthreads();
  // part-major: slot i of part p goes to [c][p * PK + i] (coalesced); the
  // combine kernel reads it back by key offset
  const size_t RP = size_t(P) << pkb;
  const size_t k0 = size_t(p) << pkb;
  for (int i = threadIdx.x; i < PK; i += RP_ANT) {
    pcnt[size_t(c) * RP + k0 + i] = cnt[i];
#pragma unroll
    for (int a = 0; a < NA; a++) pacc[(size_t(c) * NA + a) * RP + k0 + i] = acc[size_t(a) * PK + i];
  }
}

// Per key offset k < R (index rp_index(k) of part (k / 2^RP_BLK) mod P):
// the chunks' counts and accumulators combined in chunk order into entry k
// of `w` ({key bits, count, accumulators}); flag[k] = 1 when the key
// occurs.
__global__ __launch_bounds__(256) void rp_combine_kernel(
    const unsigned* __restrict__ pcnt, const unsigned long long* __restrict__ pacc, int C,
    unsigned long long R, int P, int pkb, AccSet accs, unsigned long long kmin, int stride,
    unsigned long long* __restrict__ w, unsigned* __restrict__ flag) {
  const size_t RP = size_t(P) << pkb;  // slots of one chunk's arrays
  const int ppb = __builtin_ctz(unsigned(P));
  for (size_t k = size_t(blockIdx.x) * 256 + threadIdx.x; k < R; k += size_t(gridDim.x) * 256) {
    const size_t slot = (((k >> RP_BLK) & size_t(P - 1)) << pkb) + rp_index(k, ppb);
    unsigned long long cn = 0;
    for (int c = 0; c < C; c++) cn += pcnt[size_t(c) * RP + slot];
    unsigned long long* e = w + k * size_t(stride);
    e[0] = kmin + k;
    e[1] = cn;
    for (int a = 0; a < accs.n; a++) {
      unsigned long long x = acc_identity(accs.a[a].kind, accs.a[a].vtype);
      for (int c = 0; c < C; c++)
        x = acc_combine(x, pacc[(size_t(c) * accs.n + a) * RP + slot], accs.a[a].kind,
                        accs.a[a].vtype);
      e[2 + a] = x;
    }
    flag[k] = cn != 0;
  }
}

// The heavy keys' entries ({key bits, count, accumulators}, *hot_n of them,
// a count on the device) go to their key's entry of `w`.
__global__ __launch_bounds__(256) void rp_hot_kernel(const unsigned long long* __restrict__ hot,
                                                     const unsigned* __restrict__ hot_n,
                                                     unsigned long long kmin,
                                                     int stride, unsigned long long* __restrict__ w,
                                                     unsigned* __restrict__ flag) {
  const unsigned m = *hot_n;
  for (unsigned j = blockIdx.x * 256 + threadIdx.x; j < m; j += gridDim.x * 256) {
    const unsigned long long* src = hot + size_t(j) * stride;
    const unsigned long long k = src[0] - kmin;
    for (int q = 0; q < stride; q++) w[k * stride + q] = src[q];
    flag[k] = 1;
  }
}

// ent[rank] = k for every present key offset k (incl: inclusive prefix sum
// of the flags).
__global__ __launch_bounds__(256) void rp_ent_kernel(const unsigned* __restrict__ flag,
                                                     const unsigned* __restrict__ incl,
                                                     unsigned long long R,
                                                     size_t* __restrict__ ent) {
  for (size_t k = size_t(blockIdx.x) * 256 + threadIdx.x; k < R; k += size_t(gridDim.x) * 256)
    if (flag[k]) ent[incl[k] - 1] = k;
}

// The range path after the heavy-key pass (its cold regions: nregions of
// HOT_CHUNK rows, seg_n rows each; the heavy keys' entries in hot_fin).
// kmin: the smallest key's bits; R: the key range.
// Offsets per part of the range path with na accumulators: the largest
// power of two whose arrays fit RP_LDS.
int rp_part_bits(int na) {
  const size_t bpk = 4 + 8 * size_t(na);
  int pkb = 0;
  while (pkb < 16 && (size_t(2) << pkb) * bpk <= RP_LDS) pkb++;
  return pkb;
}

// cktype: the type of the cold keys as the heavy-key pass wrote them (the
// key type, or FHIP_U32 for 64-bit keys kept as their low 32 bits)
void group_by_agg_range(int ktype, int cktype, size_t n, const AccSet& accs,
                        const fhip_agg* aggs, int naggs, const std::vector<int>& slot_of,
                        const void* ckeys, const unsigned long long* seg_h, size_t nregions,
                        const unsigned long long* hot_fin, unsigned hot_max,
                        const unsigned* hot_n, unsigned long long kmin, unsigned long long R, void* keys_out,
                        size_t capacity, size_t* ngroups, hipStream_t st, const DBuf* cold_cols) {
  const int na = accs.n;
  const int stride = 2 + na;
  const size_t bpk = 4 + 8 * size_t(na);
  const int pkb = rp_part_bits(na);
  int P = 1;  // a power of two: key offset k goes to part (k / 2^RP_BLK) mod P
  while ((size_t(P) << pkb) < size_t(R)) P *= 2;
  // the tiles of the cold regions (seg_h: each region's cold rows, on the
  // host), packed as first row << 14 | rows - 1
  unsigned long long* tl =
      static_cast<unsigned long long*>(pinned_scratch(1, nregions * RP_TPR * 8));
  size_t ntiles = 0;
  for (size_t b = 0; b < nregions; b++)
    for (unsigned long long off = 0; off < seg_h[b]; off += RP_TILE)
      tl[ntiles++] = ((b * HOT_CHUNK + off) << 14) |
                     (std::min<unsigned long long>(RP_TILE, seg_h[b] - off) - 1);
  DBuf tiles(ntiles * 8);
  FHIP_CHECK(hipMemcpyAsync(tiles.get(), tl, ntiles * 8, hipMemcpyHostToDevice, st));
  const size_t prow = ntiles * RP_TILE;  // slots of the partitioned copy
  DBuf lk(prow * sizeof(unsigned short)), starts((size_t(P) + 1) * ntiles * sizeof(unsigned short));
  DBuf pc[HA_MAX_ACC];
  Cols cold{}, pout{};
  cold.n = pout.n = na;
  AccSet paccs = accs;
  for (int a = 0; a < na; a++) {
    cold.bytes[a] = pout.bytes[a] = type_bytes(accs.a[a].vtype);
    cold.in[a] = cold_cols[a].get();
    pc[a].reset(prow * size_t(pout.bytes[a]));
    pout.out[a] = pc[a].get();
    paccs.a[a].col = pc[a].get();
  }
  {
    ProfScope prof("groupby_range_part", st);
    auto launch = [&](auto kz) {
      using KK = decltype(kz);
      const KK* kp = static_cast<const KK*>(ckeys);
      const unsigned g = unsigned(ceil_div(ntiles, size_t(8)) * 8);
      switch (na) {
        case 0: rp_part_kernel<KK, 0><<<g, RP_NT, 0, st>>>(kp, cold, tiles.as<unsigned long long>(), ntiles, kmin, P, lk.as<unsigned short>(), pout, starts.as<unsigned short>()); break;
        case 1: rp_part_kernel<KK, 1><<<g, RP_NT, 0, st>>>(kp, cold, tiles.as<unsigned long long>(), ntiles, kmin, P, lk.as<unsigned short>(), pout, starts.as<unsigned short>()); break;
        case 2: rp_part_kernel<KK, 2><<<g, RP_NT, 0, st>>>(kp, cold, tiles.as<unsigned long long>(), ntiles, kmin, P, lk.as<unsigned short>(), pout, starts.as<unsigned short>()); break;
        case 3: rp_part_kernel<KK, 3><<<g, RP_NT, 0, st>>>(kp, cold, tiles.as<unsigned long long>(), ntiles, kmin, P, lk.as<unsigned short>(), pout, starts.as<unsigned short>()); break;
        default: rp_part_kernel<KK, 4><<<g, RP_NT, 0, st>>>(kp, cold, tiles.as<unsigned long long>(), ntiles, kmin, P, lk.as<unsigned short>(), pout, starts.as<unsigned short>()); break;
      }
    };
    switch (cktype) {
      case FHIP_I32: launch(int32_t()); break;
      case FHIP_U32: launch(uint32_t()); break;
      case FHIP_I64: launch(int64_t()); break;
      default: launch(uint64_t()); break;
    }
    FHIP_CHECK(hipGetLastError());
  }
  // chunks per part: at least 512 workgroups (two a CU), each with at least
  // 64 tiles a wave; more chunks cost the combine a pass over their arrays
  const long long cdef = tune_get("groupby.range_chunks", 0);
  int C = cdef > 0 ? int(cdef) : int(std::max<size_t>(1, ceil_div(512, size_t(P))));
  C = int(std::max<size_t>(1, std::min<size_t>(size_t(C), ntiles / (64 * (RP_ANT / 64)))));
  const size_t RP = size_t(P) << pkb;  // slots of one chunk's arrays (>= R)
  DBuf pcnt(size_t(C) * RP * 4), pacc(size_t(C) * RP * 8 * size_t(std::max(na, 1)));
  {
    ProfScope prof("groupby_range_agg", st);
    const size_t lds = (size_t(1) << pkb) * bpk;
    static bool attr = false;
    if (!attr) {
      for (auto f : {(const void*)rp_agg_kernel<0>, (const void*)rp_agg_kernel<1>,
                     (const void*)rp_agg_kernel<2>, (const void*)rp_agg_kernel<3>,
                     (const void*)rp_agg_kernel<4>})
        FHIP_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, int(RP_LDS)));
      attr = true;
    }
    const unsigned g = unsigned(size_t(P) * size_t(C));
    const unsigned short* sp = starts.as<unsigned short>();
    unsigned* cp = pcnt.as<unsigned>();
    unsigned long long* ap = pacc.as<unsigned long long>();
    switch (na) {
      case 0: rp_agg_kernel<0><<<g, RP_ANT, lds, st>>>(lk.as<unsigned short>(), paccs, sp, ntiles, P, C, pkb, cp, ap); break;
      case 1: rp_agg_kernel<1><<<g, RP_ANT, lds, st>>>(lk.as<unsigned short>(), paccs, sp, ntiles, P, C, pkb, cp, ap); break;
      case 2: rp_agg_kernel<2><<<g, RP_ANT, lds, st>>>(lk.as<unsigned short>(), paccs, sp, ntiles, P, C, pkb, cp, ap); break;
      case 3: rp_agg_kernel<3><<<g, RP_ANT, lds, st>>>(lk.as<unsigned short>(), paccs, sp, ntiles, P, C, pkb, cp, ap); break;
      default: rp_agg_kernel<4><<<g, RP_ANT, lds, st>>>(lk.as<unsigned short>(), paccs, sp, ntiles, P, C, pkb, cp, ap); break;
    }
    FHIP_CHECK(hipGetLastError());
  }
  DBuf w(size_t(R) * stride * 8), flag(size_t(R) * 4), incl(size_t(R) * 4), ent(size_t(R) * sizeof(size_t));
  {
    ProfScope prof("groupby_range_emit", st);
    rp_combine_kernel<<<grid_for(size_t(R)), 256, 0, st>>>(
        pcnt.as<unsigned>(), pacc.as<unsigned long long>(), C, R, P, pkb, accs, kmin, stride,
        w.as<unsigned long long>(), flag.as<unsigned>());
    FHIP_CHECK(hipGetLastError());
    if (hot_max) {
      rp_hot_kernel<<<grid_for(hot_max), 256, 0, st>>>(hot_fin, hot_n, kmin, stride,
                                                         w.as<unsigned long long>(),
                                                         flag.as<unsigned>());
      FHIP_CHECK(hipGetLastError());
    }
    prefix_sum(FHIP_U32, flag.get(), incl.get(), size_t(R), st);
    rp_ent_kernel<<<grid_for(size_t(R)), 256, 0, st>>>(flag.as<unsigned>(), incl.as<unsigned>(), R,
                                                       ent.as<size_t>());
    FHIP_CHECK(hipGetLastError());
  }
  // the group count stays on the device (the last inclusive sum): the
  // gather and the emit take at most min(count, capacity) groups from it, so
  // the host waits once, at the end
  const unsigned* gdev = incl.as<unsigned>() + (R - 1);
  const size_t gmax = std::min<size_t>(size_t(R), capacity);
  const unsigned long long* wp = w.as<unsigned long long>();
  const size_t* ep = ent.as<size_t>();
  if (gmax > 0) {
    const int g = grid_for(gmax);
    switch (ktype) {
      case FHIP_I32: gather_group_keys<int32_t><<<g, 256, 0, st>>>(wp, stride, ep, gmax, static_cast<int32_t*>(keys_out), gdev); break;
      case FHIP_U32: gather_group_keys<uint32_t><<<g, 256, 0, st>>>(wp, stride, ep, gmax, static_cast<uint32_t*>(keys_out), gdev); break;
      case FHIP_I64: gather_group_keys<int64_t><<<g, 256, 0, st>>>(wp, stride, ep, gmax, static_cast<int64_t*>(keys_out), gdev); break;
      default: gather_group_keys<uint64_t><<<g, 256, 0, st>>>(wp, stride, ep, gmax, static_cast<uint64_t*>(keys_out), gdev); break;
    }
    FHIP_CHECK(hipGetLastError());
    if (naggs > 0) {
      EmitSet es{};
      es.n = naggs;
      for (int i = 0; i < naggs; i++) {
        es.op[i] = aggs[i].op;
        es.vtype[i] = aggs[i].vtype;
        es.slot[i] = slot_of[i] < 0 ? 0 : slot_of[i];
        es.out[i] = aggs[i].out;
      }
      emit_kernel<<<g, 256, 0, st>>>(wp, stride, ep, gmax, es, gdev);
      FHIP_CHECK(hipGetLastError());
    }
  }
  unsigned* gh = static_cast<unsigned*>(pinned_scratch(0, sizeof(unsigned)));
  FHIP_CHECK(hipMemcpyAsync(gh, gdev, sizeof(unsigned), hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  const size_t G = *gh;
  *ngroups = G;  // reported even when it does not fit, so the caller can resize
  FHIP_REQUIRE(G <= capacity, FHIP_ERR_TOO_LARGE,
               "group_by_agg: output capacity " + std::to_string(capacity) + " < groups " +
                   std::to_string(G));
}

}  // namespace

void group_by_agg(int ktype, const void* keys, size_t n, const fhip_agg* aggs, int naggs,
                  void* keys_out, size_t capacity, size_t* ngroups, size_t table_hint,
                  hipStream_t st, const HeavyKeys* heavy, bool narrow_cold) {
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
  const int kb = type_bytes(ktype);
  const int stride = 2 + accs.n;
  // LDS table slots (power of two, capped by LDS) and the expected groups per
  // partition (load factor) come from the tunables.  2048 slots by default:
  // two aggregation workgroups share a CU, which outweighs the longer runs
  // per partition that half as many partitions would give the scatter
  int S = int(tune_get("groupby.slots", PA_SLOTS));
  S = std::max(256, std::min(S, accs.n <= 2 ? 2 * PA_SLOTS : PA_SLOTS));
  while (S & (S - 1)) S &= S - 1;
  const long long fill_pct = std::max(5LL, std::min(90LL, tune_get("groupby.fill_pct", 50)));
  const unsigned long long chunk = std::max(4096LL, tune_get("groupby.chunk", (long long)PA_CHUNK));
  const int span_tiles = int(std::max(1LL, std::min(256LL, tune_get("groupby.span", PS_SPAN))));
  const size_t lds_table_bytes = size_t(S + 1) * 8 * size_t(2 + accs.n);
  static bool attr_set = false;
  if (!attr_set) {
    const int mx = int(std::max(size_t(2 * PA_SLOTS + 1) * 8 * 4,
                                size_t(PA_SLOTS + 1) * 8 * (2 + HA_MAX_ACC)));
    auto set_agg = [&](auto kz) {
      using KK = decltype(kz);
      FHIP_CHECK(hipFuncSetAttribute((const void*)pa_agg_kernel<KK, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
      FHIP_CHECK(hipFuncSetAttribute((const void*)pa_agg_kernel<KK, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
      FHIP_CHECK(hipFuncSetAttribute((const void*)pa_agg_kernel<KK, 2>, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
      FHIP_CHECK(hipFuncSetAttribute((const void*)pa_agg_kernel<KK, 3>, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
      FHIP_CHECK(hipFuncSetAttribute((const void*)pa_agg_kernel<KK, 4>, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
    };
    set_agg(int32_t());
    set_agg(uint32_t());
    set_agg(int64_t());
    set_agg(uint64_t());
    FHIP_CHECK(hipFuncSetAttribute((const void*)pa_merge_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
    const int ms = int(size_t(PS_TILE) * 10 + (size_t(1) << PA_MAX_PBITS) * 16);
    const int mw = int(size_t(PS_NT) * PS_ITEMS_WIDE * 10 + size_t(PS_WIDE_MAX_P) * 16);
    auto set_scatter = [&](auto kz) {
      using KK = decltype(kz);
      FHIP_CHECK(hipFuncSetAttribute((const void*)pa_scatter_kernel<KK, PS_ITEMS>, hipFuncAttributeMaxDynamicSharedMemorySize, ms));
      FHIP_CHECK(hipFuncSetAttribute((const void*)pa_scatter_kernel<KK, PS_ITEMS_WIDE>, hipFuncAttributeMaxDynamicSharedMemorySize, mw));
    };
    set_scatter(int32_t());
    set_scatter(uint32_t());
    set_scatter(int64_t());
    set_scatter(uint64_t());
    attr_set = true;
  }
  // partitions of the cold path: about S * fill_pct / 100 expected groups
  // each; with a table hint the count is known before the heavy-key pass,
  // which then also counts the cold rows per partition
  const size_t per_part = std::max<size_t>(1, size_t(S) * size_t(fill_pct) / 100);
  auto pbits_for = [&](size_t want) {
    int pb = 4;
    while (pb < PA_MAX_PBITS && (size_t(1) << pb) * per_part < want) pb++;
    return pb;
  };
  // ---- heavy keys: aggregated in one pass, the other rows go on ----------
  HotPlan hp;
  if (heavy) {
    hp = plan_hot_known(*heavy, n, S);
  } else {
    ProfScope prof("groupby_plan", st);
    hp = plan_hot(ktype, keys, n, S, st);
  }
  // the range path (group_by_agg_range) needs the key range, which the
  // heavy-key pass records; the hash partitions' cold-row counts are then
  // not made in that pass
  const bool range_on = hp.used && tune_get("groupby.range", 1) != 0;
  // 64-bit keys whose heavy keys lie within the range path's reach: the
  // heavy-key pass keeps the cold keys' low 32 bits only (the range path
  // needs no more: its keys span less than 2^32 values, so a key is its
  // range's smallest key plus the low bits' difference mod 2^32); if the
  // range path does not apply after all, the group-by runs again with whole
  // cold keys
  bool cold32 = false;
  if (range_on && narrow_cold && kb == 8 && tune_get("groupby.cold32", 1) != 0) {
    const unsigned long long flip =
        (ktype == FHIP_I32 || ktype == FHIP_I64) ? (1ull << 63) : 0ull;
    unsigned long long lo = ~0ull, hi = 0ull;
    for (unsigned long long kbits : hp.table) {
      if (kbits == HA_EMPTY) continue;
      lo = std::min(lo, kbits ^ flip);
      hi = std::max(hi, kbits ^ flip);
    }
    cold32 = lo <= hi && hi - lo < RP_MAXR;
  }
  const void* ckeys = keys;  // the rows the partitioned path sees
  size_t cn = n;
  AccSet caccs = accs;
  // hot_cnt: {cold rows, heavy groups}, then every region's cold rows (seg),
  // then with the range path every region's cold-key range: one readback
  DBuf hot_tab, cold_k, cold_c[HA_MAX_ACC], hot_part, hot_cnt, hot_fin, hot_ccnt;
  unsigned long long* hot_seg = nullptr;
  unsigned long long* hot_range = nullptr;
  int hot_cpbits = -1;  // partition bits of the cold-row counts made by the heavy-key pass
  // with the heavy-key pass, the cold rows of hot workgroup b sit at rows
  // b * HOT_CHUNK .. b * HOT_CHUNK + seg[b] - 1 of the cold buffers, and the
  // partitioned path takes each such region as one span
  const unsigned long long* seg_n = nullptr;
  size_t seg_spans = 0;
  unsigned hot_groups = 0;
  unsigned long long hc[2] = {0, 0};  // cold rows, heavy-key groups
  bool cn_pending = false;            // hc not read back yet
  if (hp.used) {
    const int Sh = hp.Sh;
    static bool hot_attr = false;
    if (!hot_attr) {
      // (2 + accumulators) words per slot (at most 4 x 4096 or 6 x 2048
      // words), plus the cold-row partition counts
      const int mx = HOT_MAX_LDS;
      auto set_all = [&](auto kz) {
        using KK = decltype(kz);
        FHIP_CHECK(hipFuncSetAttribute((const void*)hot_agg_kernel<KK, 0, KK>, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
        FHIP_CHECK(hipFuncSetAttribute((const void*)hot_agg_kernel<KK, 1, KK>, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
        FHIP_CHECK(hipFuncSetAttribute((const void*)hot_agg_kernel<KK, 2, KK>, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
        FHIP_CHECK(hipFuncSetAttribute((const void*)hot_agg_kernel<KK, 3, KK>, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
        FHIP_CHECK(hipFuncSetAttribute((const void*)hot_agg_kernel<KK, 4, KK>, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
        if constexpr (sizeof(KK) == 8) {
          FHIP_CHECK(hipFuncSetAttribute((const void*)hot_agg_kernel<KK, 0, uint32_t>, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
          FHIP_CHECK(hipFuncSetAttribute((const void*)hot_agg_kernel<KK, 1, uint32_t>, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
          FHIP_CHECK(hipFuncSetAttribute((const void*)hot_agg_kernel<KK, 2, uint32_t>, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
          FHIP_CHECK(hipFuncSetAttribute((const void*)hot_agg_kernel<KK, 3, uint32_t>, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
          FHIP_CHECK(hipFuncSetAttribute((const void*)hot_agg_kernel<KK, 4, uint32_t>, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
        }
      };
      set_all(int32_t());
      set_all(uint32_t());
      set_all(int64_t());
      set_all(uint64_t());
      hot_attr = true;
    }
    FHIP_REQUIRE(size_t(2 + accs.n) * Sh * 8 <= size_t(HOT_MAX_LDS), FHIP_ERR_ARG,
                 "group_by_agg: heavy-key table exceeds LDS");
    hot_tab.reset(size_t(Sh) * 8);
    // uploaded from pinned slot 3 (next used by the final assembly, after
    // several waits on the stream)
    void* tab_h = pinned_scratch(3, size_t(Sh) * 8);
    std::memcpy(tab_h, hp.table.data(), size_t(Sh) * 8);
    FHIP_CHECK(hipMemcpyAsync(hot_tab.get(), tab_h, size_t(Sh) * 8, hipMemcpyHostToDevice, st));
    cold_k.reset(n * size_t(cold32 ? 4 : kb));
    Cols cold{};
    cold.n = accs.n;
    for (int a = 0; a < accs.n; a++) {
      cold.bytes[a] = type_bytes(accs.a[a].vtype);
      cold_c[a].reset(n * size_t(cold.bytes[a]));
      cold.out[a] = cold_c[a].get();
      cold.in[a] = nullptr;
    }
    const size_t nparts = ceil_div(n, HOT_CHUNK);
    hot_part.reset(nparts * size_t(1 + accs.n) * Sh * 8);
    hot_cnt.reset(16 + nparts * 8 + (range_on ? nparts * 16 : 0));
    FHIP_CHECK(hipMemsetAsync(hot_cnt.get(), 0, 16, st));
    hot_seg = hot_cnt.as<unsigned long long>() + 2;
    if (range_on) hot_range = hot_seg + nparts;
    unsigned long long* cold_n = hot_cnt.as<unsigned long long>();
    unsigned* hot_n = reinterpret_cast<unsigned*>(cold_n + 1);
    size_t lds = size_t(2 + accs.n) * Sh * 8;
    if (table_hint && !range_on) {
      const int cp = pbits_for(table_hint);
      if (lds + (size_t(4) << cp) <= size_t(HOT_MAX_LDS)) {
        hot_cpbits = cp;
        lds += size_t(4) << cp;
        hot_ccnt.reset((size_t(8) << cp) * nparts);
      }
    }
    {
      ProfScope prof("groupby_hot", st);
      auto launch = [&](auto kz, auto cz) {
        using KK = decltype(kz);
        using CK = decltype(cz);
        const KK* kp = static_cast<const KK*>(keys);
        CK* ck = cold_k.as<CK>();
        unsigned long long* ht = hot_tab.as<unsigned long long>();
        unsigned long long* hpart = hot_part.as<unsigned long long>();
        unsigned long long* sg = hot_seg;
        unsigned long long* cc = hot_cpbits >= 0 ? hot_ccnt.as<unsigned long long>() : nullptr;
        unsigned long long* kr = hot_range;
        switch (accs.n) {
          case 0: hot_agg_kernel<KK, 0, CK><<<unsigned(nparts), HOT_NT, lds, st>>>(kp, n, accs, ht, Sh, hpart, ck, cold, cold_n, sg, hot_cpbits, cc, kr); break;
          case 1: hot_agg_kernel<KK, 1, CK><<<unsigned(nparts), HOT_NT, lds, st>>>(kp, n, accs, ht, Sh, hpart, ck, cold, cold_n, sg, hot_cpbits, cc, kr); break;
          case 2: hot_agg_kernel<KK, 2, CK><<<unsigned(nparts), HOT_NT, lds, st>>>(kp, n, accs, ht, Sh, hpart, ck, cold, cold_n, sg, hot_cpbits, cc, kr); break;
          case 3: hot_agg_kernel<KK, 3, CK><<<unsigned(nparts), HOT_NT, lds, st>>>(kp, n, accs, ht, Sh, hpart, ck, cold, cold_n, sg, hot_cpbits, cc, kr); break;
          default: hot_agg_kernel<KK, 4, CK><<<unsigned(nparts), HOT_NT, lds, st>>>(kp, n, accs, ht, Sh, hpart, ck, cold, cold_n, sg, hot_cpbits, cc, kr); break;
        }
      };
      switch (ktype) {
        case FHIP_I32: launch(int32_t(), int32_t()); break;
        case FHIP_U32: launch(uint32_t(), uint32_t()); break;
        case FHIP_I64: if (cold32) launch(int64_t(), uint32_t()); else launch(int64_t(), int64_t()); break;
        default: if (cold32) launch(uint64_t(), uint32_t()); else launch(uint64_t(), uint64_t()); break;
      }
      FHIP_CHECK(hipGetLastError());
    }
    // with the range path the host decides from the cold rows' count, every
    // region's cold rows and cold-key range, which are final after the
    // heavy-key pass: one copy, and the host waits for it (an event) while
    // the heavy tables merge.  rb[2], rb[3]: the counts; rb + 4: the regions'
    // rows; rb + 4 + nparts: their ranges
    unsigned long long* rb = nullptr;
    hipEvent_t meta_ready = nullptr;
    if (hot_cpbits < 0) {
      rb = static_cast<unsigned long long*>(pinned_scratch(0, 32 + (range_on ? nparts * 24 : 0)));
      if (range_on) {
        FHIP_CHECK(hipMemcpyAsync(rb + 2, hot_cnt.get(), 16 + nparts * 24, hipMemcpyDeviceToHost, st));
        meta_ready = thread_event();
        FHIP_CHECK(hipEventRecord(meta_ready, st));
      }
    }
    hot_fin.reset(size_t(S + 1) * stride * 8);
    // two levels: groups of 16 tables, then the groups' tables
    constexpr int per = 16;
    const int ng = int(ceil_div(nparts, size_t(per)));
    DBuf fold(size_t(ng) * size_t(1 + accs.n) * Sh * 8);
    {
      ProfScope prof("groupby_hot_merge", st);
      hot_fold_kernel<<<dim3(unsigned(Sh / 64), unsigned(ng)), 256, 0, st>>>(
          hot_part.as<unsigned long long>(), int(nparts), per, Sh, accs, fold.as<unsigned long long>());
      hot_merge_kernel<<<unsigned(Sh / 64), 256, 0, st>>>(
          fold.as<unsigned long long>(), ng, Sh, accs, hot_tab.as<unsigned long long>(),
          stride, hot_fin.as<unsigned long long>(), hot_n);
      FHIP_CHECK(hipGetLastError());
    }
    if (hot_cpbits >= 0) {
      // the partitioned path's first steps need neither: both come back
      // with the partition sizes, so the host does not wait here
      cn_pending = true;
    } else {
      // the heavy groups' count is final once the tables have merged: the
      // range path reads it on the device, the other paths wait for it
      auto read_counts = [&] {
        FHIP_CHECK(hipMemcpyAsync(rb + 2, hot_cnt.get(), 16, hipMemcpyDeviceToHost, st));
        FHIP_CHECK(hipStreamSynchronize(st));
        hot_groups = unsigned(rb[3] & 0xFFFFFFFFull);
      };
      if (range_on) FHIP_CHECK(hipEventSynchronize(meta_ready));
      else read_counts();
      hc[0] = rb[2];
      if (range_on) {  // the cold keys' range over the workgroups' ranges
        rb[2] = ~0ull;
        rb[3] = 0ull;
        for (size_t b = 0; b < nparts; b++) {
          rb[2] = std::min(rb[2], rb[4 + nparts + 2 * b]);
          rb[3] = std::max(rb[3], rb[4 + nparts + 2 * b + 1]);
        }
      }
      cn = size_t(hc[0]);
      if (range_on && cn > 0 && rb[2] <= rb[3]) {
        // the cold keys' range (from the heavy-key pass) and the heavy keys'
        const unsigned long long flip0 =
            (ktype == FHIP_I32 || ktype == FHIP_I64) ? (1ull << 63) : 0ull;
        for (unsigned long long kb : hp.table) {
          if (kb == HA_EMPTY) continue;
          rb[2] = std::min(rb[2], kb ^ flip0);
          rb[3] = std::max(rb[3], kb ^ flip0);
        }
        // keys within R consecutive values: the range path, when its parts
        // (2^pkb offsets each) number at most RP_MAXP
        const unsigned long long R = rb[3] - rb[2] + 1;
        const unsigned long long rmax =
            (unsigned long long)std::max(0LL, tune_get("groupby.range_max", (long long)RP_MAXR));
        const int pkb = rp_part_bits(accs.n);
        if (R != 0 && R <= n && R <= std::min(rmax, RP_MAXR) &&
            size_t(R) <= (size_t(RP_MAXP) << pkb)) {
          const unsigned long long flip =
              (ktype == FHIP_I32 || ktype == FHIP_I64) ? (1ull << 63) : 0ull;
          group_by_agg_range(ktype, cold32 ? FHIP_U32 : ktype, n, accs, aggs, naggs, slot_of, cold_k.get(),
                             rb + 4, nparts, hot_fin.as<unsigned long long>(), unsigned(Sh),
                             hot_n, rb[2] ^ flip, R, keys_out, capacity, ngroups, st, cold_c);
          return;
        }
      }
      if (range_on) read_counts();
      hc[1] = hot_groups;
      if (cold32 && cn > 0) {
        // the cold keys were kept as their low 32 bits for a range path that
        // does not apply: again, with whole cold keys
        group_by_agg(ktype, keys, n, aggs, naggs, keys_out, capacity, ngroups, table_hint, st,
                     heavy, false);
        return;
      }
    }
    ckeys = cold_k.get();
    for (int a = 0; a < accs.n; a++) caccs.a[a].col = cold_c[a].get();
    seg_n = hot_seg;
    seg_spans = nparts;
  }

  // Final assembly over Ptot partitions' tables (the heavy keys are the
  // last one when used): compact, sort the group keys, gather the outputs.
  // low_bytes: the key bytes the group-key sort must run (-1: unknown)
  // range (may be null): the smallest and largest group key, as ordered key
  // bits (sign bit flipped for signed keys)
  auto assemble = [&](const unsigned long long* w, const unsigned* fn_dev,
                      const std::vector<unsigned>& fn, int Ptot, int low_bytes,
                      const unsigned long long* range) {
    // partition offsets, uploaded from pinned memory (the lambda ends with
    // a synchronisation, so the slot is free again on return)
    unsigned long long* foff =
        static_cast<unsigned long long*>(pinned_scratch(3, size_t(Ptot) * 8));
    size_t G = 0;
    for (int p = 0; p < Ptot; p++) {
      foff[p] = G;
      G += fn[p];
    }
    *ngroups = G;  // reported even when it does not fit, so the caller can resize
    FHIP_REQUIRE(G <= capacity, FHIP_ERR_TOO_LARGE,
                 "group_by_agg: output capacity " + std::to_string(capacity) + " < groups " +
                     std::to_string(G));
    if (G == 0) return;
    DBuf fob(size_t(Ptot) * 8), entb(G * sizeof(size_t));
    FHIP_CHECK(hipMemcpyAsync(fob.get(), foff, size_t(Ptot) * 8, hipMemcpyHostToDevice, st));
    final_index_kernel<<<Ptot, 256, 0, st>>>(fn_dev, fob.as<unsigned long long>(), Ptot, S,
                                              entb.as<size_t>());
    FHIP_CHECK(hipGetLastError());
    auto gather_keys = [&](const size_t* ent, void* dst) {
      switch (ktype) {
        case FHIP_I32: gather_group_keys<int32_t><<<grid_for(G), 256, 0, st>>>(w, stride, ent, G, static_cast<int32_t*>(dst)); break;
        case FHIP_U32: gather_group_keys<uint32_t><<<grid_for(G), 256, 0, st>>>(w, stride, ent, G, static_cast<uint32_t*>(dst)); break;
        case FHIP_I64: gather_group_keys<int64_t><<<grid_for(G), 256, 0, st>>>(w, stride, ent, G, static_cast<int64_t*>(dst)); break;
        default: gather_group_keys<uint64_t><<<grid_for(G), 256, 0, st>>>(w, stride, ent, G, static_cast<uint64_t*>(dst)); break;
      }
      FHIP_CHECK(hipGetLastError());
    };
    // keys are distinct, so ascending order is the reference's order: when
    // they fill their range densely enough (at most `groupby.dense_ratio`
    // possible keys per group, 32 by default; at most 2^28), a key's rank is
    // the number of group keys at or below it (flags and a scan); else the
    // group keys are radix-sorted
    const long long ratio = tune_get("groupby.dense_ratio", 32);
    const unsigned long long R = range ? range[1] - range[0] + 1 : 0;
    DBuf ent_sorted;
    const size_t* ent_out = entb.as<size_t>();
    if (range && range[0] <= range[1] && R != 0 && ratio > 0 && R <= (size_t(1) << 28) &&
        R <= (unsigned long long)(ratio) * G) {
      const unsigned long long mask = kb == 8 ? ~0ull : (1ull << (8 * kb)) - 1;
      const unsigned long long flip =
          (ktype == FHIP_I32 || ktype == FHIP_I64) ? (1ull << (8 * kb - 1)) : 0ull;
      DBuf flag(size_t(R) * 4), incl(size_t(R) * 4);
      ent_sorted.reset(G * sizeof(size_t));
      FHIP_CHECK(hipMemsetAsync(flag.get(), 0, size_t(R) * 4, st));
      dense_mark_kernel<<<grid_for(G), 256, 0, st>>>(w, stride, entb.as<size_t>(), G, mask, flip,
                                                     range[0], flag.as<unsigned>());
      FHIP_CHECK(hipGetLastError());
      prefix_sum(FHIP_U32, flag.get(), incl.get(), size_t(R), st);
      dense_order_kernel<<<grid_for(G), 256, 0, st>>>(w, stride, entb.as<size_t>(), G, mask,
                                                      flip, range[0], incl.as<unsigned>(),
                                                      ent_sorted.as<size_t>());
      FHIP_CHECK(hipGetLastError());
      ent_out = ent_sorted.as<size_t>();
      gather_keys(ent_out, keys_out);
    } else {
      DBuf gk(G * size_t(kb));
      gather_keys(entb.as<size_t>(), gk.get());
      sort_pairs(ktype, gk.get(), gk.get(), 8, entb.get(), entb.get(), G, false, false, st,
                 low_bytes);
      FHIP_CHECK(hipMemcpyAsync(keys_out, gk.get(), G * size_t(kb), hipMemcpyDeviceToDevice, st));
    }
    if (naggs > 0) {
      EmitSet es{};
      es.n = naggs;
      for (int i = 0; i < naggs; i++) {
        es.op[i] = aggs[i].op;
        es.vtype[i] = aggs[i].vtype;
        es.slot[i] = slot_of[i] < 0 ? 0 : slot_of[i];
        es.out[i] = aggs[i].out;
      }
      emit_kernel<<<grid_for(G), 256, 0, st>>>(w, stride, ent_out, G, es);
      FHIP_CHECK(hipGetLastError());
    }
    FHIP_CHECK(hipStreamSynchronize(st));  // the pinned slot of foff is reused by the next call
  };

  auto assemble_hot_only = [&]() {  // every row had a heavy key
    DBuf fnb(sizeof(unsigned));
    FHIP_CHECK(hipMemcpyAsync(fnb.get(), &hot_groups, sizeof(unsigned), hipMemcpyHostToDevice, st));
    assemble(hot_fin.as<unsigned long long>(), fnb.as<unsigned>(), std::vector<unsigned>{hot_groups}, 1, -1,
             nullptr);
  };
  if (hp.used && !cn_pending && cn == 0) {
    assemble_hot_only();
    return;
  }

  int pbits = pbits_for(table_hint ? table_hint : std::min<size_t>(cn, size_t(1) << 20));
  for (;;) {
    const int P = 1 << pbits;
    const int Ptot = P + (hp.used ? 1 : 0);
    const size_t tiles = seg_n ? seg_spans : ceil_div(ceil_div(cn, PS_TILE), size_t(span_tiles));  // spans
    // the heavy-key pass counted the cold rows for this partition count
    const bool counted = seg_n && pbits == hot_cpbits;
    DBuf cb(counted ? 0 : size_t(P) * tiles * 8), ib(size_t(P) * tiles * 8), ovf(sizeof(int));
    unsigned long long* cnt = counted ? hot_ccnt.as<unsigned long long>() : cb.as<unsigned long long>();
    unsigned long long* incl = ib.as<unsigned long long>();
    FHIP_CHECK(hipMemsetAsync(ovf.get(), 0, sizeof(int), st));
    DBuf pkeys;  // the partitioned copy, allocated once the cold-row count is known
    DBuf pcols[HA_MAX_ACC];
    Cols cols{};
    cols.n = caccs.n;
    for (int a = 0; a < caccs.n; a++) {
      cols.bytes[a] = type_bytes(caccs.a[a].vtype);
      cols.in[a] = caccs.a[a].col;
    }
    auto run = [&](int pass) {
      switch (ktype) {
        case FHIP_I32: launch_pa<int32_t>(ckeys, cn, pbits, tiles, span_tiles, cols, cnt, incl, pkeys.get(), st, pass, seg_n, HOT_CHUNK); break;
        case FHIP_U32: launch_pa<uint32_t>(ckeys, cn, pbits, tiles, span_tiles, cols, cnt, incl, pkeys.get(), st, pass, seg_n, HOT_CHUNK); break;
        case FHIP_I64: launch_pa<int64_t>(ckeys, cn, pbits, tiles, span_tiles, cols, cnt, incl, pkeys.get(), st, pass, seg_n, HOT_CHUNK); break;
        default: launch_pa<uint64_t>(ckeys, cn, pbits, tiles, span_tiles, cols, cnt, incl, pkeys.get(), st, pass, seg_n, HOT_CHUNK); break;
      }
    };
    {
      ProfScope prof("groupby_part_hist", st);
      if (!counted) run(1);
      prefix_sum(FHIP_U64, cnt, incl, size_t(P) * tiles, st);
    }
    // partition sizes (inclusive offset at the last span of every partition)
    // and, while pending, the cold-row count come back through pinned
    // memory; the scatter is queued before the host waits for them
    unsigned long long* ends =
        static_cast<unsigned long long*>(pinned_scratch(0, (size_t(P) + 2) * 8));
    DBuf eb(size_t(P) * 8);
    gather_strided_u64_kernel<<<grid_for_n(size_t(P)), 256, 0, st>>>(
        incl + (tiles - 1), tiles, size_t(P), eb.as<unsigned long long>());
    FHIP_CHECK(hipGetLastError());
    FHIP_CHECK(hipMemcpyAsync(ends, eb.get(), size_t(P) * 8, hipMemcpyDeviceToHost, st));
    if (cn_pending)
      FHIP_CHECK(hipMemcpyAsync(ends + P, hot_cnt.get(), 16, hipMemcpyDeviceToHost, st));
    hipEvent_t sizes_ready = thread_event();
    FHIP_CHECK(hipEventRecord(sizes_ready, st));
    // the partitioned copy holds the cold rows (every row, while their count
    // is still on its way: the cold buffers' own bound)
    const size_t prows = cn_pending ? n : cn;
    pkeys.reset(prows * size_t(kb));
    for (int a = 0; a < caccs.n; a++) {
      pcols[a].reset(prows * size_t(cols.bytes[a]));
      cols.out[a] = pcols[a].get();
    }
    {
      ProfScope prof("groupby_part_scatter", st);
      run(2);
    }
    FHIP_CHECK(hipEventSynchronize(sizes_ready));
    if (cn_pending) {
      cn_pending = false;
      cn = size_t(ends[P]);
      hot_groups = unsigned(ends[P + 1] & 0xFFFFFFFFull);
      if (cn == 0) {
        assemble_hot_only();
        return;
      }
    }
    // chunks of at most `chunk` rows per partition, built while the scatter
    // runs, in pinned memory (items, then the first item of every partition)
    size_t NI = 0;
    {
      unsigned long long prev = 0;
      for (int p = 0; p < P; p++) {
        const unsigned long long e = ends[p];
        NI += e == prev ? 1 : size_t(ceil_div(size_t(e - prev), size_t(chunk)));
        prev = e;
      }
    }
    const size_t items_bytes = (NI * sizeof(Item) + 15) & ~size_t(15);
    char* stage = static_cast<char*>(pinned_scratch(1, items_bytes + (size_t(P) + 1) * sizeof(unsigned)));
    Item* items = reinterpret_cast<Item*>(stage);
    unsigned* part_items = reinterpret_cast<unsigned*>(stage + items_bytes);
    {
      size_t ni = 0;
      unsigned long long run_off = 0;
      for (int p = 0; p < P; p++) {
        part_items[p] = unsigned(ni);
        const unsigned long long e = ends[p];
        for (unsigned long long b = run_off; b < e; b += chunk)
          items[ni++] = Item{b, std::min<unsigned long long>(b + chunk, e)};
        if (e == run_off) items[ni++] = Item{run_off, run_off};
        run_off = e;
      }
      part_items[P] = unsigned(ni);
    }
    DBuf itb(NI * sizeof(Item)), pib((P + 1) * sizeof(unsigned));
    FHIP_CHECK(hipMemcpyAsync(itb.get(), items, NI * sizeof(Item), hipMemcpyHostToDevice, st));
    FHIP_CHECK(hipMemcpyAsync(pib.get(), part_items, (P + 1) * sizeof(unsigned),
                              hipMemcpyHostToDevice, st));
    DBuf partial(NI * size_t(S + 1) * stride * 8), pn(NI * sizeof(unsigned));
    FHIP_CHECK(hipMemsetAsync(pn.get(), 0, NI * sizeof(unsigned), st));
    AccSet paccs = caccs;  // the partitioned copies of the columns
    for (int a = 0; a < caccs.n; a++) paccs.a[a].col = pcols[a].get();
    {
      ProfScope prof("hashagg", st);
      auto launch_agg = [&](auto kz) {
        using KK = decltype(kz);
        const KK* pk = pkeys.as<KK>();
        unsigned long long* pp = partial.as<unsigned long long>();
        switch (paccs.n) {
          case 0: pa_agg_kernel<KK, 0><<<unsigned(NI), PA_NT, lds_table_bytes, st>>>(itb.as<Item>(), pk, paccs, S, stride, pp, pn.as<unsigned>(), ovf.as<int>()); break;
          case 1: pa_agg_kernel<KK, 1><<<unsigned(NI), PA_NT, lds_table_bytes, st>>>(itb.as<Item>(), pk, paccs, S, stride, pp, pn.as<unsigned>(), ovf.as<int>()); break;
          case 2: pa_agg_kernel<KK, 2><<<unsigned(NI), PA_NT, lds_table_bytes, st>>>(itb.as<Item>(), pk, paccs, S, stride, pp, pn.as<unsigned>(), ovf.as<int>()); break;
          case 3: pa_agg_kernel<KK, 3><<<unsigned(NI), PA_NT, lds_table_bytes, st>>>(itb.as<Item>(), pk, paccs, S, stride, pp, pn.as<unsigned>(), ovf.as<int>()); break;
          default: pa_agg_kernel<KK, 4><<<unsigned(NI), PA_NT, lds_table_bytes, st>>>(itb.as<Item>(), pk, paccs, S, stride, pp, pn.as<unsigned>(), ovf.as<int>()); break;
        }
      };
      switch (ktype) {
        case FHIP_I32: launch_agg(int32_t()); break;
        case FHIP_U32: launch_agg(uint32_t()); break;
        case FHIP_I64: launch_agg(int64_t()); break;
        default: launch_agg(uint64_t()); break;
      }
      FHIP_CHECK(hipGetLastError());
    }
    DBuf fin(size_t(Ptot) * (S + 1) * stride * 8), fnb(size_t(Ptot) * sizeof(unsigned));
    {
      ProfScope prof("groupby_merge", st);
      pa_merge_kernel<<<P, PA_NT, lds_table_bytes, st>>>(
          pib.as<unsigned>(), partial.as<unsigned long long>(), pn.as<unsigned>(), caccs, S, stride,
          fin.as<unsigned long long>(), fnb.as<unsigned>(), ovf.as<int>());
      FHIP_CHECK(hipGetLastError());
    }
    if (hp.used) {  // the heavy keys' table is the last partition
      FHIP_CHECK(hipMemcpyAsync(fin.as<unsigned long long>() + size_t(P) * (S + 1) * stride,
                                hot_fin.get(), size_t(hot_groups) * stride * 8,
                                hipMemcpyDeviceToDevice, st));
      FHIP_CHECK(hipMemcpyAsync(fnb.as<unsigned>() + P, &hot_groups, sizeof(unsigned),
                                hipMemcpyHostToDevice, st));
    }
    // the key range of all the final tables (read back with the counts)
    DBuf rng(2 * sizeof(unsigned long long));
    FHIP_CHECK(hipMemsetAsync(rng.get(), 0xFF, sizeof(unsigned long long), st));
    FHIP_CHECK(hipMemsetAsync(rng.as<unsigned long long>() + 1, 0, sizeof(unsigned long long), st));
    key_range_kernel<<<unsigned(Ptot), 256, 0, st>>>(
        fin.as<unsigned long long>(), stride, S, fnb.as<unsigned>(), kb,
        (ktype == FHIP_I32 || ktype == FHIP_I64) ? 1 : 0, rng.as<unsigned long long>());
    FHIP_CHECK(hipGetLastError());
    // the overflow flag, the group count of every partition, then the range
    const size_t rb_words = (size_t(P) + 2) & ~size_t(1);  // u32 words before the range
    unsigned* rb = static_cast<unsigned*>(pinned_scratch(2, rb_words * 4 + 16));
    unsigned long long* rng_h = reinterpret_cast<unsigned long long*>(rb + rb_words);
    FHIP_CHECK(hipMemcpyAsync(rb, ovf.get(), sizeof(int), hipMemcpyDeviceToHost, st));
    FHIP_CHECK(hipMemcpyAsync(rb + 1, fnb.get(), size_t(P) * sizeof(unsigned),
                              hipMemcpyDeviceToHost, st));
    FHIP_CHECK(hipMemcpyAsync(rng_h, rng.get(), 16, hipMemcpyDeviceToHost, st));
    FHIP_CHECK(hipStreamSynchronize(st));
    const int ov = int(rb[0]);
    std::vector<unsigned> fn(rb + 1, rb + 1 + P);
    fn.resize(Ptot);
    if (ov) {  // a partition had more groups than its LDS table: more partitions
      if (pbits < PA_MAX_PBITS) {
        pbits = std::min(pbits + 2, PA_MAX_PBITS);
        continue;
      }
      group_by_agg_sorted(ktype, keys, n, aggs, naggs, keys_out, capacity, ngroups, st);
      return;
    }
    if (hp.used) fn[P] = hot_groups;
    // the bytes below the highest one where the smallest and largest keys differ
    int low_bytes = -1;
    if (rng_h[0] <= rng_h[1]) {
      const unsigned long long d = rng_h[0] ^ rng_h[1];
      low_bytes = d ? (63 - __builtin_clzll(d)) / 8 + 1 : 0;
    }
    assemble(fin.as<unsigned long long>(), fnb.as<unsigned>(), fn, Ptot, low_bytes,
             rng_h[0] <= rng_h[1] ? rng_h : nullptr);
    return;
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
