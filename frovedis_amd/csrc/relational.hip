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
//     columns, and the group keys of every column.
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

// Run boundary flags of one column over the current row order: flag[i] = 1
// where col[idx[i]] != col[idx[i - 1]] (the comparison of set_separate,
// core/set_operations.hpp:146-163).  `first` overwrites, later columns OR.
template <class T>
__global__ __launch_bounds__(RL_NT) void boundary_kernel(const T* __restrict__ col,
                                                         const size_t* __restrict__ idx, size_t n,
                                                         uint8_t* __restrict__ flag, int first) {
  const size_t stride = size_t(gridDim.x) * RL_NT;
  for (size_t i = size_t(blockIdx.x) * RL_NT + threadIdx.x; i < n; i += stride) {
    const bool b = i > 0 && col[idx[i]] != col[idx[i - 1]];
    if (first) flag[i] = b ? 1 : 0;
    else if (b) flag[i] = 1;
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

void multi_group_by(int ncols, const int* ktypes, const void* const* cols, size_t* idx, size_t n,
                    size_t* split, size_t* nsplit, void* const* keys_out, hipStream_t st) {
  FHIP_REQUIRE(ncols >= 1, FHIP_ERR_ARG, "multi_group_by: no key column");
  for (int c = 0; c < ncols; c++) type_bytes(ktypes[c]);
  if (n == 0) {
    *nsplit = set_separate(ktypes[0], nullptr, 0, split, st);
    return;
  }
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
  // (set_separate per column, merged by set_union)
  {
    ProfScope prof("multi_group_split", st);
    for (int c = 0; c < ncols; c++) {
      dispatch_type(ktypes[c], [&](auto z) {
        using T = decltype(z);
        boundary_kernel<T><<<grid_for_n(n), RL_NT, 0, st>>>((const T*)cols[c], idx, n,
                                                            flag.as<uint8_t>(), c == 0);
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
