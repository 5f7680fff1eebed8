// group_by and grouped aggregation on one rank's share, gfx950.
//
// group_by follows dataframe/dfcolumn_groupby.cc:16-105 with one node:
// gather the key column through local_idx, stable radix sort of
// (key, row index), set_separate, first key of each run.  The outputs are
// exactly the reference's local_grouped_idx, local_idx_split and group keys.
//
// Grouped sum/count/avg/max/min follow dataframe/dfcolumn_grouped_agg.cc
// (x86 bodies, merge with one node).  Each group is reduced in row order by
// contiguous blocks: one wave per group up to kBig rows, 8192-row chunks
// for larger groups whose partials are folded in chunk order.  Integer sums
// wrap in T exactly like the reference's sequential fold; FP sums are
// deterministic but use a tree order, so they match the reference's fold to
// rounding.  max/min keep the first of equal values, as the sequential
// `if (m < v) m = v` does.
#include "fhip_device.hpp"
#include "radix_sort.hpp"

#include <limits>

namespace fhip {

size_t set_separate(int type, const void* sorted, size_t n, size_t* out, hipStream_t st);
void prefix_sum(int type, const void* in, void* out, size_t n, hipStream_t st);

namespace {

constexpr size_t kBig = 8192;  // rows handled by one wave
constexpr int AG_NT = 256;
constexpr int AG_PER_THREAD = int(kBig / AG_NT);

enum { OP_SUM = 0, OP_COUNT = 1, OP_AVG = 2, OP_MAX = 3, OP_MIN = 4 };

template <class T, int OP>
struct AggOp {
  // value a null row reads as (sum: 0, max: lowest(), min: max())
  __host__ __device__ static T null_value() {
    if (OP == OP_MAX) return std::numeric_limits<T>::lowest();
    if (OP == OP_MIN) return std::numeric_limits<T>::max();
    return T(0);
  }
  __host__ __device__ static T identity() { return null_value(); }
  __device__ static T fold(T a, T v) {
    if (OP == OP_MAX) return (a < v) ? v : a;
    if (OP == OP_MIN) return (a > v) ? v : a;
    return a + v;
  }
  // left is earlier in row order
  __device__ static T combine(T l, T r) { return fold(l, r); }
  __device__ static T finish(T a) {
    if (OP == OP_SUM) return T(0) + a;  // 0 + x: no -0.0 sums
    return a;
  }
};

// COUNT is a sum of 1 per non-null row, carried as uint64
struct CountOp {
  __device__ static unsigned long long identity() { return 0; }
  __device__ static unsigned long long combine(unsigned long long l, unsigned long long r) { return l + r; }
};

template <class T, int OP>
__device__ __forceinline__ T load_val(const T* col, const size_t* gidx, const uint8_t* nullmask,
                                      size_t i) {
  size_t r = gidx[i];
  if (nullmask && nullmask[r]) return AggOp<T, OP>::null_value();
  return col[r];
}

// in-order tree combine across the 64 lanes (lane 0 ends with the result)
template <class T, class C>
__device__ __forceinline__ T wave_combine_ordered(T acc, C comb) {
  const unsigned lane = __lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    T other = __shfl_down(acc, off, 64);
    if ((lane & (2 * off - 1)) == 0) acc = comb(acc, other);
  }
  return acc;
}

template <class T, int OP>
__global__ __launch_bounds__(256) void agg_small_kernel(const T* __restrict__ col,
                                                        const size_t* __restrict__ gidx,
                                                        const size_t* __restrict__ split, size_t G,
                                                        const uint8_t* __restrict__ nullmask,
                                                        T* __restrict__ out) {
  using A = AggOp<T, OP>;
  const unsigned lane = __lane_id();
  const size_t nw = size_t(gridDim.x) * (blockDim.x / 64);
  for (size_t g = (size_t(blockIdx.x) * blockDim.x + threadIdx.x) / 64; g < G; g += nw) {
    const size_t s = split[g], e = split[g + 1], len = e - s;
    if (len > kBig) continue;
    const size_t per = (len + 63) / 64;
    const size_t lo = s + lane * per;
    const size_t hi = lo + per < e ? lo + per : e;
    T acc = A::identity();
    for (size_t i = lo; i < hi; i++) acc = A::fold(acc, load_val<T, OP>(col, gidx, nullmask, i));
    acc = wave_combine_ordered(acc, [](T l, T r) { return A::combine(l, r); });
    if (lane == 0) out[g] = A::finish(acc);
  }
}

__global__ __launch_bounds__(256) void count_small_kernel(const size_t* __restrict__ gidx,
                                                          const size_t* __restrict__ split,
                                                          size_t G,
                                                          const uint8_t* __restrict__ nullmask,
                                                          unsigned long long* __restrict__ out) {
  const unsigned lane = __lane_id();
  const size_t nw = size_t(gridDim.x) * (blockDim.x / 64);
  for (size_t g = (size_t(blockIdx.x) * blockDim.x + threadIdx.x) / 64; g < G; g += nw) {
    const size_t s = split[g], e = split[g + 1], len = e - s;
    if (!nullmask) {  // size_helper: split differences (:719-745)
      if (lane == 0) out[g] = len;
      continue;
    }
    if (len > kBig) continue;
    unsigned long long c = 0;
    for (size_t i = s + lane; i < e; i += 64) c += nullmask[gidx[i]] ? 0 : 1;
    c = wave_sum(c);
    if (lane == 0) out[g] = c;
  }
}

// chunks of big groups: nch_incl is the inclusive scan of chunks per group
__device__ __forceinline__ size_t chunk_group(const unsigned long long* nch_incl, size_t G,
                                              size_t k) {
  size_t lo = 0, hi = G;  // first g with nch_incl[g] > k
  while (lo < hi) {
    size_t mid = (lo + hi) >> 1;
    if (nch_incl[mid] <= k) lo = mid + 1; else hi = mid;
  }
  return lo;
}

template <class T, int OP>
__global__ __launch_bounds__(AG_NT) void agg_chunk_kernel(
    const T* __restrict__ col, const size_t* __restrict__ gidx, const size_t* __restrict__ split,
    size_t G, const uint8_t* __restrict__ nullmask, const unsigned long long* __restrict__ nch_incl,
    T* __restrict__ partial) {
  using A = AggOp<T, OP>;
  const size_t k = blockIdx.x;
  const size_t g = chunk_group(nch_incl, G, k);
  const size_t nch = nch_incl[g] - (g ? nch_incl[g - 1] : 0);
  const size_t c = k - (nch_incl[g] - nch);
  const size_t s = split[g] + c * kBig;
  const size_t e = split[g + 1] < s + kBig ? split[g + 1] : s + kBig;
  const size_t lo = s + size_t(threadIdx.x) * AG_PER_THREAD;
  const size_t hi = lo + AG_PER_THREAD < e ? lo + AG_PER_THREAD : e;
  T acc = A::identity();
  for (size_t i = lo; i < hi; i++) acc = A::fold(acc, load_val<T, OP>(col, gidx, nullmask, i));
  acc = wave_combine_ordered(acc, [](T l, T r) { return A::combine(l, r); });
  __shared__ T ws[AG_NT / 64];
  if (__lane_id() == 0) ws[threadIdx.x / 64] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    T r = ws[0];
    for (int w = 1; w < AG_NT / 64; w++) r = A::combine(r, ws[w]);
    partial[k] = r;
  }
}

__global__ __launch_bounds__(AG_NT) void count_chunk_kernel(
    const size_t* __restrict__ gidx, const size_t* __restrict__ split, size_t G,
    const uint8_t* __restrict__ nullmask, const unsigned long long* __restrict__ nch_incl,
    unsigned long long* __restrict__ partial) {
  const size_t k = blockIdx.x;
  const size_t g = chunk_group(nch_incl, G, k);
  const size_t nch = nch_incl[g] - (g ? nch_incl[g - 1] : 0);
  const size_t c = k - (nch_incl[g] - nch);
  const size_t s = split[g] + c * kBig;
  const size_t e = split[g + 1] < s + kBig ? split[g + 1] : s + kBig;
  unsigned long long cnt = 0;
  for (size_t i = s + threadIdx.x; i < e; i += AG_NT) cnt += nullmask[gidx[i]] ? 0 : 1;
  cnt = wave_sum(cnt);
  __shared__ unsigned long long ws[AG_NT / 64];
  if (__lane_id() == 0) ws[threadIdx.x / 64] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) partial[k] = ws[0] + ws[1] + ws[2] + ws[3];
}

template <class T, class C, class F>
__global__ void combine_chunks_kernel(const unsigned long long* __restrict__ nch_incl, size_t G,
                                      const T* __restrict__ partial, T ident, C comb, F fin,
                                      T* __restrict__ out) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x; g < G; g += stride) {
    const size_t end = nch_incl[g];
    const size_t beg = g ? nch_incl[g - 1] : 0;
    if (beg == end) continue;
    T acc = ident;
    for (size_t k = beg; k < end; k++) acc = comb(acc, partial[k]);
    out[g] = fin(acc);
  }
}

__global__ void nchunks_kernel(const size_t* __restrict__ split, size_t G,
                               unsigned long long* __restrict__ nch) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x; g < G; g += stride) {
    size_t len = split[g + 1] - split[g];
    nch[g] = len > kBig ? (len + kBig - 1) / kBig : 0;
  }
}

__global__ void nullmask_kernel(const size_t* __restrict__ nulls, size_t nn, size_t n,
                                uint8_t* __restrict__ mask) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nn; i += stride)
    if (nulls[i] < n) mask[nulls[i]] = 1;
}

template <class T>
__global__ void null_fix_kernel(T* __restrict__ out, const unsigned long long* __restrict__ cnt,
                                size_t G, int op) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x; g < G; g += stride) {
    if (op == OP_SUM && cnt[g] == 0) out[g] = std::numeric_limits<T>::max();
    if (op == OP_MAX && out[g] == std::numeric_limits<T>::lowest())
      out[g] = std::numeric_limits<T>::max();
  }
}

template <class T>
__global__ void avg_kernel(const T* __restrict__ sum, const unsigned long long* __restrict__ cnt,
                           size_t G, int has_nulls, double* __restrict__ out) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x; g < G; g += stride) {
    if (has_nulls && cnt[g] == 0) out[g] = double(std::numeric_limits<T>::max());
    else out[g] = static_cast<double>(sum[g]) / static_cast<double>(cnt[g]);
  }
}

template <class K>
__global__ void gather_kernel(const K* __restrict__ col, const size_t* __restrict__ idx, size_t n,
                              K* __restrict__ out) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = col[idx[i]];
}

template <class K>
__global__ void group_keys_kernel(const K* __restrict__ sorted, const size_t* __restrict__ split,
                                  size_t G, K* __restrict__ keys) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x; g < G; g += stride)
    keys[g] = sorted[split[g]];
}

int grid_for(size_t n, int per_cu = 8) {
  return int(std::max<size_t>(1, std::min<size_t>(ceil_div(n, 256), size_t(fill_blocks(256, per_cu)))));
}

template <class F>
void dispatch(int t, F&& f) {
  switch (t) {
    case FHIP_I32: f(int32_t()); break;
    case FHIP_U32: f(uint32_t()); break;
    case FHIP_I64: f(int64_t()); break;
    case FHIP_U64: f(uint64_t()); break;
    case FHIP_F32: f(float()); break;
    case FHIP_F64: f(double()); break;
    default: throw Error(FHIP_ERR_TYPE, "unsupported type code " + std::to_string(t));
  }
}

// Chunk bookkeeping for groups larger than kBig; returns total chunks.
size_t plan_chunks(const size_t* split, size_t G, DBuf& nch, hipStream_t st) {
  nch.reset((G + 1) * sizeof(unsigned long long));
  DBuf raw((G + 1) * sizeof(unsigned long long));
  nchunks_kernel<<<grid_for(G), 256, 0, st>>>(split, G, raw.as<unsigned long long>());
  FHIP_CHECK(hipGetLastError());
  prefix_sum(FHIP_U64, raw.get(), nch.get(), G, st);
  unsigned long long total = 0;
  FHIP_CHECK(hipMemcpyAsync(&total, nch.as<unsigned long long>() + (G - 1), sizeof(total),
                            hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  return size_t(total);
}

template <class T, int OP>
void agg_values(const T* col, const size_t* gidx, const size_t* split, size_t G,
                const uint8_t* mask, T* out, hipStream_t st) {
  using A = AggOp<T, OP>;
  agg_small_kernel<T, OP><<<grid_for(G * 64), 256, 0, st>>>(col, gidx, split, G, mask, out);
  FHIP_CHECK(hipGetLastError());
  DBuf nch;
  size_t K = plan_chunks(split, G, nch, st);
  if (K == 0) return;
  DBuf part(K * sizeof(T));
  agg_chunk_kernel<T, OP><<<unsigned(K), AG_NT, 0, st>>>(col, gidx, split, G, mask,
                                                         nch.as<unsigned long long>(), part.as<T>());
  FHIP_CHECK(hipGetLastError());
  combine_chunks_kernel<<<grid_for(G), 256, 0, st>>>(
      nch.as<unsigned long long>(), G, part.as<T>(), A::identity(),
      [] __device__(T l, T r) { return A::combine(l, r); },
      [] __device__(T a) { return A::finish(a); }, out);
  FHIP_CHECK(hipGetLastError());
}

void count_values(const size_t* gidx, const size_t* split, size_t G, const uint8_t* mask,
                  unsigned long long* out, hipStream_t st) {
  count_small_kernel<<<grid_for(G * 64), 256, 0, st>>>(gidx, split, G, mask, out);
  FHIP_CHECK(hipGetLastError());
  if (!mask) return;
  DBuf nch;
  size_t K = plan_chunks(split, G, nch, st);
  if (K == 0) return;
  DBuf part(K * sizeof(unsigned long long));
  count_chunk_kernel<<<unsigned(K), AG_NT, 0, st>>>(gidx, split, G, mask,
                                                    nch.as<unsigned long long>(),
                                                    part.as<unsigned long long>());
  FHIP_CHECK(hipGetLastError());
  combine_chunks_kernel<<<grid_for(G), 256, 0, st>>>(
      nch.as<unsigned long long>(), G, part.as<unsigned long long>(), 0ull,
      [] __device__(unsigned long long l, unsigned long long r) { return l + r; },
      [] __device__(unsigned long long a) { return a; }, out);
  FHIP_CHECK(hipGetLastError());
}

}  // namespace

void group_by(int ktype, const void* col, size_t* idx, size_t n, size_t* split, size_t* nsplit,
              void* keys_out, hipStream_t st) {
  const int kb = type_bytes(ktype);
  if (n == 0) {
    *nsplit = set_separate(ktype, nullptr, 0, split, st);
    return;
  }
  DBuf nv(n * size_t(kb));
  dispatch(ktype, [&](auto z) {
    using K = decltype(z);
    gather_kernel<K><<<grid_for(n), 256, 0, st>>>((const K*)col, idx, n, nv.as<K>());
  });
  FHIP_CHECK(hipGetLastError());
  sort_pairs(ktype, nv.get(), nv.get(), 8, idx, idx, n, false, false, st);
  *nsplit = set_separate(ktype, nv.get(), n, split, st);
  const size_t G = *nsplit - 1;
  dispatch(ktype, [&](auto z) {
    using K = decltype(z);
    group_keys_kernel<K><<<grid_for(G), 256, 0, st>>>(nv.as<K>(), split, G, (K*)keys_out);
  });
  FHIP_CHECK(hipGetLastError());
}

void grouped_agg(int op, int vtype, const void* col, size_t col_n, const size_t* gidx, size_t n,
                 const size_t* split, size_t nsplit, const size_t* nulls, size_t nn, void* out,
                 hipStream_t st) {
  FHIP_REQUIRE(op >= OP_SUM && op <= OP_MIN, FHIP_ERR_ARG, "grouped_agg: bad op");
  if (nsplit < 2) return;
  const size_t G = nsplit - 1;
  DBuf maskb;
  const uint8_t* mask = nullptr;
  if (nn) {  // one flag per row of the column (grouped_idx indexes the column)
    maskb.reset(col_n);
    FHIP_CHECK(hipMemsetAsync(maskb.get(), 0, col_n, st));
    nullmask_kernel<<<grid_for(nn), 256, 0, st>>>(nulls, nn, col_n, maskb.as<uint8_t>());
    FHIP_CHECK(hipGetLastError());
    mask = maskb.as<uint8_t>();
  }
  if (op == OP_COUNT) {
    count_values(gidx, split, G, mask, (unsigned long long*)out, st);
    return;
  }
  dispatch(vtype, [&](auto z) {
    using T = decltype(z);
    if (op == OP_MAX) {
      agg_values<T, OP_MAX>((const T*)col, gidx, split, G, mask, (T*)out, st);
      if (mask) {
        null_fix_kernel<T><<<grid_for(G), 256, 0, st>>>((T*)out, nullptr, G, OP_MAX);
        FHIP_CHECK(hipGetLastError());
      }
    } else if (op == OP_MIN) {
      agg_values<T, OP_MIN>((const T*)col, gidx, split, G, mask, (T*)out, st);
    } else {
      DBuf sumb, cntb;
      T* sum = (T*)out;
      if (op == OP_AVG) {
        sumb.reset(G * sizeof(T));
        sum = sumb.as<T>();
      }
      agg_values<T, OP_SUM>((const T*)col, gidx, split, G, mask, sum, st);
      if (mask || op == OP_AVG) {
        cntb.reset(G * sizeof(unsigned long long));
        count_values(gidx, split, G, mask, cntb.as<unsigned long long>(), st);
      }
      if (op == OP_SUM && mask) {
        null_fix_kernel<T><<<grid_for(G), 256, 0, st>>>(sum, cntb.as<unsigned long long>(), G, OP_SUM);
        FHIP_CHECK(hipGetLastError());
      }
      if (op == OP_AVG) {
        avg_kernel<T><<<grid_for(G), 256, 0, st>>>(sum, cntb.as<unsigned long long>(), G,
                                                   mask ? 1 : 0, (double*)out);
        FHIP_CHECK(hipGetLastError());
      }
    }
  });
}

}  // namespace fhip

extern "C" {

int fhip_dev_group_by(int ktype, const void* col, size_t* idx, size_t n, size_t* split,
                      size_t* nsplit, void* keys_out, void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(split && nsplit, FHIP_ERR_ARG, "group_by: null output");
    FHIP_REQUIRE(n == 0 || (col && idx && keys_out), FHIP_ERR_ARG, "group_by: null pointer");
    fhip::group_by(ktype, col, idx, n, split, nsplit, keys_out, fhip::as_stream(stream));
  });
}

int fhip_dev_grouped_agg(int op, int vtype, const void* col, size_t col_n, const size_t* gidx,
                         size_t n, const size_t* split, size_t nsplit, const size_t* nulls,
                         size_t nnull, void* out, void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(nsplit < 2 || (split && out && gidx && (col || op == 1)), FHIP_ERR_ARG,
                 "grouped_agg: null pointer");
    FHIP_REQUIRE(nnull == 0 || nulls, FHIP_ERR_ARG, "grouped_agg: null nulls list");
    fhip::grouped_agg(op, vtype, col, col_n, gidx, n, split, nsplit, nulls, nnull, out,
                      fhip::as_stream(stream));
  });
}

}  // extern "C"
