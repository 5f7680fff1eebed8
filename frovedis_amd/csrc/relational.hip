// The dataframe operations on either side of the execution core, one rank's
// share, for gfx950:
//   - gather (global_extract on one rank, dfcolumn_misc.cc:714-761, then
//     append_nulls :1052-1078): the rows of a column at given indices, NULL
//     (numeric_limits<T>::max(), the reference's NULL value) where the
//     index is FHIP_NULL_INDEX -- the right-side columns of an outer join's
//     left-only rows -- and the null list of the result
//     (global_extract_null_helper2, dfcolumn.cc:223-233);
//   - filter against a scalar or another column (find_condition and
//     find_condition_pair, core/find_condition.hpp:23-152,412-546, with the
//     predicates of dfcolumn_filter_{eq,lt,le,gt,ge}.cc): the ascending row
//     positions where the comparison holds, NULL rows removed (the
//     reference's set_difference with the null lists);
//   - sort_index (dfcolumn_sort.cc:69-88,141-180 and sort_with_idx
//     :299-340 on one rank): the stable order of a column, ascending
//     (radix_sort) or descending (radix_sort_desc), and the sorted values;
//   - multi_group_by (dftable.cc:367-487 with one rank ->
//     dfcolumn_groupby.cc:122-234): stable sort of the row index by every
//     key column, last column first, the union of the run boundaries of all
//     columns, and the group keys of every column.  When every key column
//     is an integer column and their value ranges fit in 64 bits together,
//     the columns are packed into one order-preserving composite key
//     (column 0 in the high bits) and sorted once: a stable sort by the
//     composite is the same permutation as the stable sorts column by
//     column, and two rows differ in the composite exactly when they differ
//     in some column, so the splits are the composite's run boundaries.
//     This replaces the per-column random gathers with one coalesced pass.
// Comparisons between a column of type T and a scalar or column of type U
// use C++'s usual arithmetic conversions, as the reference's templated
// predicates do.  Every kernel is a grid-stride loop over rows: these are
// streaming passes bounded by HBM bandwidth.
#include "fhip_device.hpp"
#include "radix_sort.hpp"

#include <limits>

namespace fhip {

size_t set_separate(int type, const void* sorted, size_t n, size_t* out, hipStream_t st);

namespace {

constexpr int RL_NT = 256;

template <class T>
__global__ __launch_bounds__(RL_NT) void gather_null_kernel(const T* __restrict__ src,
                                                            size_t src_n,
                                                            const size_t* __restrict__ idx,
                                                            size_t n, T* __restrict__ dst,
                                                            int* __restrict__ bad) {
  const size_t stride = size_t(gridDim.x) * RL_NT;
  for (size_t i = size_t(blockIdx.x) * RL_NT + threadIdx.x; i < n; i += stride) {
    const size_t j = idx[i];
    T v = std::numeric_limits<T>::max();
    if (j != FHIP_NULL_INDEX) {
      if (j < src_n) {
        v = src[j];
      } else {
        v = T(0);
        *bad = 1;
      }
    }
    dst[i] = v;
  }
}

// flag[r] = 1 for every listed row r < n
__global__ __launch_bounds__(RL_NT) void null_flags_kernel(const size_t* __restrict__ rows,
                                                           size_t nr, size_t n,
                                                           uint8_t* __restrict__ flag) {
  const size_t stride = size_t(gridDim.x) * RL_NT;
  for (size_t i = size_t(blockIdx.x) * RL_NT + threadIdx.x; i < nr; i += stride)
    if (rows[i] < n) flag[rows[i]] = 1;
}

struct GatherNullPred {
  const size_t* idx;
  const uint8_t* src_null;  // may be null: the source has no null rows
  size_t src_n;
  __device__ bool operator()(size_t i) const {
    const size_t j = idx[i];
    if (j == FHIP_NULL_INDEX) return true;
    return src_null && j < src_n && src_null[j];
  }
};

template <class A, class B>
__device__ __forceinline__ bool compare(A a, B b, int op) {
  switch (op) {
    case FHIP_CMP_EQ: return a == b;
    case FHIP_CMP_NEQ: return a != b;
    case FHIP_CMP_LT: return a < b;
    case FHIP_CMP_LE: return a <= b;
    case FHIP_CMP_GT: return a > b;
    default: return a >= b;
  }
}

template <class T, class U>
struct ScalarPred {
  const T* col;
  U c;
  int op;
  const uint8_t* null;
  __device__ bool operator()(size_t i) const {
    if (null && null[i]) return false;
    return compare(col[i], c, op);
  }
};

template <class T, class U>
struct PairPred {
  const T* a;
  const U* b;
  int op;
  const uint8_t* null;
  __device__ bool operator()(size_t i) const {
    if (null && null[i]) return false;
    return compare(a[i], b[i], op);
  }
};

__global__ __launch_bounds__(RL_NT) void iota_kernel(size_t* __restrict__ out, size_t n) {
  const size_t stride = size_t(gridDim.x) * RL_NT;
  for (size_t i = size_t(blockIdx.x) * RL_NT + threadIdx.x; i < n; i += stride) out[i] = i;
}

template <class T>
__global__ __launch_bounds__(RL_NT) void gather_kernel(const T* __restrict__ src,
                                                       const size_t* __restrict__ idx, size_t n,
                                                       T* __restrict__ dst) {
  const size_t stride = size_t(gridDim.x) * RL_NT;
  for (size_t i = size_t(blockIdx.x) * RL_NT + threadIdx.x; i < n; i += stride)
    dst[i] = src[idx[i]];
}

// Run boundary flags of one column already in the final row order: flag[i]
// = 1 where v[i] != v[i - 1] (the comparison of set_separate,
// core/set_operations.hpp:146-163).  `first` overwrites, later columns OR.
template <class T>
__global__ __launch_bounds__(RL_NT) void boundary_kernel(const T* __restrict__ v, size_t n,
                                                         uint8_t* __restrict__ flag, int first) {
  const size_t stride = size_t(gridDim.x) * RL_NT;
  for (size_t i = size_t(blockIdx.x) * RL_NT + threadIdx.x; i < n; i += stride) {
    const bool b = i > 0 && v[i] != v[i - 1];
    if (first) flag[i] = b ? 1 : 0;
    else if (b) flag[i] = 1;
  }
}

// ---- composite keys for integer key columns --------------------------------
constexpr int MG_MAXC = 8;

// Order-preserving map of an integer value into uint64 (signed values get
// their sign bit flipped after sign extension).
template <class T>
__host__ __device__ __forceinline__ unsigned long long to_ord(T v) {
  if constexpr (std::is_signed<T>::value)
    return static_cast<unsigned long long>(static_cast<long long>(v)) ^ (1ull << 63);
  else
    return static_cast<unsigned long long>(v);
}

template <class T>
__host__ __device__ __forceinline__ T from_ord(unsigned long long u) {
  if constexpr (std::is_signed<T>::value)
    return static_cast<T>(static_cast<long long>(u ^ (1ull << 63)));
  else
    return static_cast<T>(u);
}

struct MgCols {
  const void* col[MG_MAXC];
  int type[MG_MAXC];
  unsigned long long lo[MG_MAXC];
  int shift[MG_MAXC];
  int ncols;
};

__device__ __forceinline__ unsigned long long load_ord(const void* p, int type, size_t j) {
  switch (type) {
    case FHIP_I32: return to_ord(static_cast<const int32_t*>(p)[j]);
    case FHIP_U32: return to_ord(static_cast<const uint32_t*>(p)[j]);
    case FHIP_I64: return to_ord(static_cast<const int64_t*>(p)[j]);
    default: return to_ord(static_cast<const uint64_t*>(p)[j]);
  }
}

// [min, max] of the ordered values col[idx[i]]; mm = {min, max}, preset to
// {~0, 0}
template <class T>
__global__ __launch_bounds__(RL_NT) void range_kernel(const T* __restrict__ col,
                                                      const size_t* __restrict__ idx, size_t n,
                                                      unsigned long long* __restrict__ mm) {
  unsigned long long lo = ~0ull, hi = 0;
  const size_t stride = size_t(gridDim.x) * RL_NT;
  for (size_t i = size_t(blockIdx.x) * RL_NT + threadIdx.x; i < n; i += stride) {
    const unsigned long long u = to_ord(col[idx[i]]);
    lo = u < lo ? u : lo;
    hi = u > hi ? u : hi;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long a = __shfl_xor(lo, o, 64), b = __shfl_xor(hi, o, 64);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  __shared__ unsigned long long s_lo[RL_NT / 64], s_hi[RL_NT / 64];
  if (__lane_id() == 0) {
    s_lo[threadIdx.x / 64] = lo;
    s_hi[threadIdx.x / 64] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < RL_NT / 64; w++) {
      lo = s_lo[w] < lo ? s_lo[w] : lo;
      hi = s_hi[w] > hi ? s_hi[w] : hi;
    }
    atomicMin(mm, lo);
    atomicMax(mm + 1, hi);
  }
}

__global__ void range_init_kernel(unsigned long long* mm, int ncols) {
  const int c = threadIdx.x;
  if (c < ncols) {
    mm[2 * c] = ~0ull;
    mm[2 * c + 1] = 0;
  }
}

// comp[i] = sum over columns of (ord(col_c[idx[i]]) - lo_c) << shift_c
template <class K>
__global__ __launch_bounds__(RL_NT) void composite_kernel(MgCols c, const size_t* __restrict__ idx,
                                                          size_t n, K* __restrict__ comp) {
  const size_t stride = size_t(gridDim.x) * RL_NT;
  for (size_t i = size_t(blockIdx.x) * RL_NT + threadIdx.x; i < n; i += stride) {
    const size_t j = idx[i];
    unsigned long long key = 0;
    for (int k = 0; k < c.ncols; k++)
      key |= (load_ord(c.col[k], c.type[k], j) - c.lo[k]) << c.shift[k];
    comp[i] = static_cast<K>(key);
  }
}

template <class K>
struct CompRunStart {
  const K* k;
  __device__ bool operator()(size_t i) const { return i == 0 || k[i] != k[i - 1]; }
};

// keys[g] of one column from the composite key at the group's first row
template <class T, class K>
__global__ __launch_bounds__(RL_NT) void decode_kernel(const K* __restrict__ comp,
                                                       const size_t* __restrict__ split, size_t G,
                                                       unsigned long long lo, int shift,
                                                       unsigned long long mask,
                                                       T* __restrict__ keys) {
  const size_t stride = size_t(gridDim.x) * RL_NT;
  for (size_t g = size_t(blockIdx.x) * RL_NT + threadIdx.x; g < G; g += stride) {
    const unsigned long long u = (static_cast<unsigned long long>(comp[split[g]]) >> shift) & mask;
    keys[g] = from_ord<T>(u + lo);
  }
}

struct BoundaryPred {
  const uint8_t* flag;
  __device__ bool operator()(size_t i) const { return i == 0 || flag[i] != 0; }
};

__global__ void put_end_kernel(size_t* __restrict__ out, size_t pos, size_t n) { out[pos] = n; }

template <class T>
__global__ __launch_bounds__(RL_NT) void group_keys_kernel(const T* __restrict__ col,
                                                           const size_t* __restrict__ idx,
                                                           const size_t* __restrict__ split,
                                                           size_t G, T* __restrict__ keys) {
  const size_t stride = size_t(gridDim.x) * RL_NT;
  for (size_t g = size_t(blockIdx.x) * RL_NT + threadIdx.x; g < G; g += stride)
    keys[g] = col[idx[split[g]]];
}

// Byte flags over n rows marking the rows of up to two lists (empty when
// both lists are empty).
DBuf null_flags(const size_t* a, size_t na, const size_t* b, size_t nb, size_t n, hipStream_t st) {
  DBuf f;
  if (na + nb == 0 || n == 0) return f;
  f.reset(n);
  FHIP_CHECK(hipMemsetAsync(f.get(), 0, n, st));
  if (na) null_flags_kernel<<<grid_for_n(na), RL_NT, 0, st>>>(a, na, n, f.as<uint8_t>());
  if (nb) null_flags_kernel<<<grid_for_n(nb), RL_NT, 0, st>>>(b, nb, n, f.as<uint8_t>());
  FHIP_CHECK(hipGetLastError());
  return f;
}

void check_op(int op) {
  FHIP_REQUIRE(op >= FHIP_CMP_EQ && op <= FHIP_CMP_GE, FHIP_ERR_ARG,
               "filter: unknown comparison " + std::to_string(op));
}

}  // namespace

void gather(int type, const void* src, size_t src_n, const size_t* idx, size_t n, void* dst,
            hipStream_t st) {
  if (n == 0) return;
  DBuf bad(sizeof(int));
  FHIP_CHECK(hipMemsetAsync(bad.get(), 0, sizeof(int), st));
  {
    ProfScope prof("gather", st);
    dispatch_type(type, [&](auto z) {
      using T = decltype(z);
      gather_null_kernel<T><<<grid_for_n(n), RL_NT, 0, st>>>((const T*)src, src_n, idx, n,
                                                             (T*)dst, bad.as<int>());
    });
    FHIP_CHECK(hipGetLastError());
  }
  int b = 0;
  FHIP_CHECK(hipMemcpyAsync(&b, bad.get(), sizeof(int), hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  FHIP_REQUIRE(!b, FHIP_ERR_ARG, "gather: row index out of range");
}

size_t gather_nulls(const size_t* idx, size_t n, size_t src_n, const size_t* src_nulls,
                    size_t nnull, size_t* out, hipStream_t st) {
  if (n == 0) return 0;
  DBuf f = null_flags(src_nulls, nnull, nullptr, 0, src_n, st);
  return compact_indices(GatherNullPred{idx, nnull ? f.as<uint8_t>() : nullptr, src_n}, n, out,
                         st);
}

size_t filter_scalar(int ctype, const void* col, size_t n, int op, int stype, const void* scalar,
                     const size_t* nulls, size_t nn, size_t* out, hipStream_t st) {
  check_op(op);
  type_bytes(stype);
  if (n == 0) return 0;
  DBuf nf = null_flags(nulls, nn, nullptr, 0, n, st);
  const uint8_t* null = nn ? nf.as<uint8_t>() : nullptr;
  size_t r = 0;
  ProfScope prof("filter", st);
  dispatch_type(ctype, [&](auto zc) {
    using T = decltype(zc);
    dispatch_type(stype, [&](auto zs) {
      using U = decltype(zs);
      U c;
      std::memcpy(&c, scalar, sizeof(U));
      r = compact_indices(ScalarPred<T, U>{(const T*)col, c, op, null}, n, out, st);
    });
  });
  return r;
}

size_t filter_pair(int atype, const void* a, int btype, const void* b, size_t n, int op,
                   const size_t* a_nulls, size_t na, const size_t* b_nulls, size_t nb,
                   size_t* out, hipStream_t st) {
  check_op(op);
  if (n == 0) return 0;
  // eq skips the right column's nulls (dfcolumn_filter_eq.cc:60-102); the
  // other comparisons remove both sides' nulls
  if (op == FHIP_CMP_EQ) nb = 0;
  DBuf nf = null_flags(a_nulls, na, b_nulls, nb, n, st);
  const uint8_t* null = (na + nb) ? nf.as<uint8_t>() : nullptr;
  size_t r = 0;
  ProfScope prof("filter", st);
  dispatch_type(atype, [&](auto za) {
    using T = decltype(za);
    dispatch_type(btype, [&](auto zb) {
      using U = decltype(zb);
      r = compact_indices(PairPred<T, U>{(const T*)a, (const U*)b, op, null}, n, out, st);
    });
  });
  return r;
}

void sort_index(int ktype, const void* col, size_t col_n, const size_t* idx_in, size_t n,
                bool desc, void* sorted_out, size_t* idx_out, hipStream_t st) {
  if (n == 0) return;
  if (idx_in) {
    dispatch_type(ktype, [&](auto z) {
      using T = decltype(z);
      gather_kernel<T><<<grid_for_n(n), RL_NT, 0, st>>>((const T*)col, idx_in, n, (T*)sorted_out);
    });
    FHIP_CHECK(hipGetLastError());
    if (idx_out != idx_in)
      FHIP_CHECK(hipMemcpyAsync(idx_out, idx_in, n * sizeof(size_t), hipMemcpyDeviceToDevice, st));
  } else {
    FHIP_REQUIRE(n == col_n, FHIP_ERR_ARG, "sort_index: n must equal col_n without idx_in");
    FHIP_CHECK(hipMemcpyAsync(sorted_out, col, n * size_t(type_bytes(ktype)),
                              hipMemcpyDeviceToDevice, st));
    iota_kernel<<<grid_for_n(n), RL_NT, 0, st>>>(idx_out, n);
    FHIP_CHECK(hipGetLastError());
  }
  sort_pairs(ktype, sorted_out, sorted_out, 8, idx_out, idx_out, n, false, desc, st);
}

namespace {

// The composite-key path; returns false (and does nothing visible) when a
// key column is not an integer column or the ranges need more than 64 bits.
bool multi_group_by_composite(int ncols, const int* ktypes, const void* const* cols, size_t* idx,
                              size_t n, size_t* split, size_t* nsplit, void* const* keys_out,
                              hipStream_t st) {
  if (ncols > MG_MAXC) return false;
  for (int c = 0; c < ncols; c++)
    if (ktypes[c] == FHIP_F32 || ktypes[c] == FHIP_F64) return false;
  DBuf mm(2 * sizeof(unsigned long long) * size_t(ncols));
  unsigned long long* dmm = mm.as<unsigned long long>();
  range_init_kernel<<<1, 64, 0, st>>>(dmm, ncols);
  FHIP_CHECK(hipGetLastError());
  {
    ProfScope prof("multi_group_range", st);
    for (int c = 0; c < ncols; c++) {
      dispatch_type(ktypes[c], [&](auto z) {
        using T = decltype(z);
        if constexpr (std::is_integral<T>::value)
          range_kernel<T><<<grid_for_n(n), RL_NT, 0, st>>>((const T*)cols[c], idx, n,
                                                           dmm + 2 * c);
      });
      FHIP_CHECK(hipGetLastError());
    }
  }
  unsigned long long hmm[2 * MG_MAXC];
  FHIP_CHECK(hipMemcpyAsync(hmm, dmm, 2 * sizeof(unsigned long long) * size_t(ncols),
                            hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  int bits[MG_MAXC], total = 0;
  for (int c = 0; c < ncols; c++) {
    const unsigned long long span = hmm[2 * c + 1] - hmm[2 * c];
    bits[c] = span ? 64 - __builtin_clzll(span) : 0;
    total += bits[c];
  }
  if (total > 64) return false;
  MgCols mc{};
  mc.ncols = ncols;
  int sh = 0;
  for (int c = ncols - 1; c >= 0; c--) {
    mc.col[c] = cols[c];
    mc.type[c] = ktypes[c];
    mc.lo[c] = hmm[2 * c];
    mc.shift[c] = bits[c] ? sh : 0;
    sh += bits[c];
  }
  auto run = [&](auto kz) {
    using K = decltype(kz);
    DBuf comp(n * sizeof(K));
    {
      ProfScope prof("multi_group_composite", st);
      composite_kernel<K><<<grid_for_n(n), RL_NT, 0, st>>>(mc, idx, n, comp.as<K>());
      FHIP_CHECK(hipGetLastError());
    }
    // one stable sort by the composite = the column-by-column stable sorts
    sort_pairs_bits(int(sizeof(K)), comp.get(), comp.get(), 8, idx, idx, n, std::max(total, 1),
                    st);
    const size_t G = compact_indices(CompRunStart<K>{comp.as<K>()}, n, split, st);
    put_end_kernel<<<1, 1, 0, st>>>(split, G, n);
    FHIP_CHECK(hipGetLastError());
    *nsplit = G + 1;
    for (int c = 0; c < ncols; c++) {
      const unsigned long long mask = bits[c] == 64 ? ~0ull : ((1ull << bits[c]) - 1);
      dispatch_type(ktypes[c], [&](auto z) {
        using T = decltype(z);
        if constexpr (std::is_integral<T>::value)
          decode_kernel<T, K><<<grid_for_n(G), RL_NT, 0, st>>>(comp.as<K>(), split, G, mc.lo[c],
                                                               mc.shift[c], mask, (T*)keys_out[c]);
      });
      FHIP_CHECK(hipGetLastError());
    }
  };
  if (total <= 32) run(uint32_t());
  else run(uint64_t());
  return true;
}

}  // namespace

void multi_group_by(int ncols, const int* ktypes, const void* const* cols, size_t* idx, size_t n,
                    size_t* split, size_t* nsplit, void* const* keys_out, hipStream_t st) {
  FHIP_REQUIRE(ncols >= 1, FHIP_ERR_ARG, "multi_group_by: no key column");
  for (int c = 0; c < ncols; c++) type_bytes(ktypes[c]);
  if (n == 0) {
    *nsplit = set_separate(ktypes[0], nullptr, 0, split, st);
    return;
  }
  if (tune_get("multi_group_by.composite", 1) &&
      multi_group_by_composite(ncols, ktypes, cols, idx, n, split, nsplit, keys_out, st))
    return;
  int kbmax = 4;
  for (int c = 0; c < ncols; c++) kbmax = std::max(kbmax, type_bytes(ktypes[c]));
  DBuf tmp(n * size_t(kbmax)), flag(n);
  // stable sort of (col[idx], idx) by every column, the last first and
  // column 0 last (multi_group_by_sort, then multi_group_by_sort_split)
  for (int c = ncols - 1; c >= 0; c--) {
    dispatch_type(ktypes[c], [&](auto z) {
      using T = decltype(z);
      gather_kernel<T><<<grid_for_n(n), RL_NT, 0, st>>>((const T*)cols[c], idx, n, tmp.as<T>());
    });
    FHIP_CHECK(hipGetLastError());
    sort_pairs(ktypes[c], tmp.get(), tmp.get(), 8, idx, idx, n, false, false, st);
  }
  // union of the run boundaries of every column in the final order
  // (set_separate per column, merged by set_union); tmp holds column 0
  // sorted, the other columns are gathered into it once each
  {
    ProfScope prof("multi_group_split", st);
    for (int c = 0; c < ncols; c++) {
      dispatch_type(ktypes[c], [&](auto z) {
        using T = decltype(z);
        if (c > 0)
          gather_kernel<T><<<grid_for_n(n), RL_NT, 0, st>>>((const T*)cols[c], idx, n,
                                                            tmp.as<T>());
        boundary_kernel<T><<<grid_for_n(n), RL_NT, 0, st>>>(tmp.as<T>(), n, flag.as<uint8_t>(),
                                                            c == 0);
      });
      FHIP_CHECK(hipGetLastError());
    }
  }
  const size_t G = compact_indices(BoundaryPred{flag.as<uint8_t>()}, n, split, st);
  put_end_kernel<<<1, 1, 0, st>>>(split, G, n);
  FHIP_CHECK(hipGetLastError());
  *nsplit = G + 1;
  // group keys of every column (multi_group_by_extract)
  for (int c = 0; c < ncols; c++) {
    dispatch_type(ktypes[c], [&](auto z) {
      using T = decltype(z);
      group_keys_kernel<T><<<grid_for_n(G), RL_NT, 0, st>>>((const T*)cols[c], idx, split, G,
                                                            (T*)keys_out[c]);
    });
    FHIP_CHECK(hipGetLastError());
  }
}

}  // namespace fhip

extern "C" {

int fhip_dev_gather(int type, const void* src, size_t src_n, const size_t* idx, size_t n,
                    void* dst, void* stream) {
  return fhip::guarded([&] {
    fhip::type_bytes(type);
    FHIP_REQUIRE(n == 0 || (idx && dst && (src || src_n == 0)), FHIP_ERR_ARG,
                 "gather: null pointer");
    fhip::gather(type, src, src_n, idx, n, dst, fhip::as_stream(stream));
  });
}

int fhip_dev_gather_nulls(const size_t* idx, size_t n, size_t src_n, const size_t* src_nulls,
                          size_t nnull, size_t* out, size_t* nout, void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(nout && (n == 0 || (idx && out)), FHIP_ERR_ARG, "gather_nulls: null pointer");
    FHIP_REQUIRE(nnull == 0 || src_nulls, FHIP_ERR_ARG, "gather_nulls: null nulls list");
    *nout = fhip::gather_nulls(idx, n, src_n, src_nulls, nnull, out, fhip::as_stream(stream));
  });
}

int fhip_dev_filter(int ctype, const void* col, size_t n, int op, int stype, const void* scalar,
                    const size_t* nulls, size_t nnull, size_t* out, size_t* nout, void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(nout && scalar && (n == 0 || (col && out)), FHIP_ERR_ARG,
                 "filter: null pointer");
    FHIP_REQUIRE(nnull == 0 || nulls, FHIP_ERR_ARG, "filter: null nulls list");
    *nout = fhip::filter_scalar(ctype, col, n, op, stype, scalar, nulls, nnull, out,
                                fhip::as_stream(stream));
  });
}

int fhip_dev_filter_pair(int atype, const void* a, int btype, const void* b, size_t n, int op,
                         const size_t* a_nulls, size_t na_null, const size_t* b_nulls,
                         size_t nb_null, size_t* out, size_t* nout, void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(nout && (n == 0 || (a && b && out)), FHIP_ERR_ARG, "filter: null pointer");
    FHIP_REQUIRE((na_null == 0 || a_nulls) && (nb_null == 0 || b_nulls), FHIP_ERR_ARG,
                 "filter: null nulls list");
    *nout = fhip::filter_pair(atype, a, btype, b, n, op, a_nulls, na_null, b_nulls, nb_null, out,
                              fhip::as_stream(stream));
  });
}

int fhip_dev_sort_index(int ktype, const void* col, size_t col_n, const size_t* idx_in, size_t n,
                        int descending, void* sorted_out, size_t* idx_out, void* stream) {
  return fhip::guarded([&] {
    fhip::type_bytes(ktype);
    FHIP_REQUIRE(n == 0 || (col && sorted_out && idx_out), FHIP_ERR_ARG,
                 "sort_index: null pointer");
    fhip::sort_index(ktype, col, col_n, idx_in, n, descending != 0, sorted_out, idx_out,
                     fhip::as_stream(stream));
  });
}

int fhip_dev_multi_group_by(int ncols, const int* ktypes, const void* const* cols, size_t* idx,
                            size_t n, size_t* split, size_t* nsplit, void* const* keys_out,
                            void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(ncols >= 1 && ktypes && cols && keys_out && split && nsplit, FHIP_ERR_ARG,
                 "multi_group_by: null pointer");
    FHIP_REQUIRE(n == 0 || idx, FHIP_ERR_ARG, "multi_group_by: null index");
    fhip::multi_group_by(ncols, ktypes, cols, idx, n, split, nsplit, keys_out,
                         fhip::as_stream(stream));
  });
}

}  // extern "C"
