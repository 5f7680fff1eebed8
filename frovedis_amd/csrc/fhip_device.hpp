// Reusable device building blocks: stable index compaction and the
// single-workgroup scan of per-tile counts.  Header-only so every kernel
// file can instantiate them with its own predicates.
#pragma once

#include "fhip_common.hpp"

namespace fhip {

constexpr int CP_NT = 256;
constexpr int CP_ITEMS = 8;
constexpr int CP_TILE = CP_NT * CP_ITEMS;

// In-place exclusive scan of `counts[0..nb)`, one workgroup; writes the
// grand total to *total.
template <class T>
__global__ __launch_bounds__(1024) void scan_counts_kernel(T* counts, size_t nb, T* total) {
  __shared__ T tmp[16];
  __shared__ T carry_s;
  if (threadIdx.x == 0) carry_s = 0;
  __syncthreads();
  for (size_t base = 0; base < nb; base += 1024) {
    size_t i = base + threadIdx.x;
    T v = i < nb ? counts[i] : T(0);
    T all;
    T e = block_excl_scan<1024>(v, tmp, &all);
    T carry = carry_s;
    if (i < nb) counts[i] = carry + e;
    __syncthreads();
    if (threadIdx.x == 0) carry_s = carry + all;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry_s;
}

template <class Pred>
__global__ __launch_bounds__(CP_NT) void compact_count_kernel(Pred pred, size_t n,
                                                              unsigned long long* counts) {
  const size_t base = size_t(blockIdx.x) * CP_TILE;
  unsigned c = 0;
#pragma unroll
  for (int j = 0; j < CP_ITEMS; j++) {
    size_t i = base + size_t(j) * CP_NT + threadIdx.x;
    if (i < n && pred(i)) c++;
  }
  c = wave_sum(c);
  __shared__ unsigned ws[CP_NT / 64];
  if (__lane_id() == 0) ws[threadIdx.x / 64] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned t = 0;
    for (int w = 0; w < CP_NT / 64; w++) t += ws[w];
    counts[blockIdx.x] = t;
  }
}

// Writes map(i) for every i with pred(i), in increasing i.
template <class Pred, class Map, class OutT>
__global__ __launch_bounds__(CP_NT) void compact_write_kernel(Pred pred, Map map, size_t n,
                                                              const unsigned long long* offs,
                                                              OutT* out) {
  __shared__ unsigned tmp[CP_NT / 64];
  const size_t base = size_t(blockIdx.x) * CP_TILE;
  unsigned long long run = offs[blockIdx.x];
#pragma unroll
  for (int j = 0; j < CP_ITEMS; j++) {
    size_t i = base + size_t(j) * CP_NT + threadIdx.x;
    unsigned f = (i < n && pred(i)) ? 1u : 0u;
    unsigned all;
    unsigned e = block_excl_scan<CP_NT>(f, tmp, &all);
    if (f) out[run + e] = map(i);
    run += all;
  }
}

struct IdentityMap {
  __device__ size_t operator()(size_t i) const { return i; }
};

// Stable compaction; returns the number of selected indices (synchronises
// the stream).  `out` must hold every selected entry.
template <class Pred, class Map, class OutT>
size_t compact(Pred pred, Map map, size_t n, OutT* out, hipStream_t st) {
  if (n == 0) return 0;
  const size_t nb = ceil_div(n, CP_TILE);
  DBuf cb((nb + 1) * sizeof(unsigned long long));
  unsigned long long* counts = cb.as<unsigned long long>();
  compact_count_kernel<<<unsigned(nb), CP_NT, 0, st>>>(pred, n, counts);
  FHIP_CHECK(hipGetLastError());
  scan_counts_kernel<<<1, 1024, 0, st>>>(counts, nb, counts + nb);
  FHIP_CHECK(hipGetLastError());
  compact_write_kernel<<<unsigned(nb), CP_NT, 0, st>>>(pred, map, n, counts, out);
  FHIP_CHECK(hipGetLastError());
  unsigned long long total = 0;
  FHIP_CHECK(hipMemcpyAsync(&total, counts + nb, sizeof(total), hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  return size_t(total);
}

template <class Pred>
size_t compact_indices(Pred pred, size_t n, size_t* out, hipStream_t st) {
  return compact(pred, IdentityMap{}, n, out, st);
}

// Calls writer(pos, i) for every i with pred(i), pos = rank among the
// selected indices (stable).
template <class Pred, class Writer>
__global__ __launch_bounds__(CP_NT) void compact_apply_kernel(Pred pred, Writer writer, size_t n,
                                                              const unsigned long long* offs) {
  __shared__ unsigned tmp[CP_NT / 64];
  const size_t base = size_t(blockIdx.x) * CP_TILE;
  unsigned long long run = offs[blockIdx.x];
#pragma unroll
  for (int j = 0; j < CP_ITEMS; j++) {
    size_t i = base + size_t(j) * CP_NT + threadIdx.x;
    unsigned f = (i < n && pred(i)) ? 1u : 0u;
    unsigned all;
    unsigned e = block_excl_scan<CP_NT>(f, tmp, &all);
    if (f) writer(size_t(run + e), i);
    run += all;
  }
}

// Counts the selected indices (synchronises); if the count is at most
// `capacity`, also runs the writer.  Returns the count.
template <class Pred, class Writer>
size_t compact_apply(Pred pred, Writer writer, size_t n, size_t capacity, hipStream_t st) {
  if (n == 0) return 0;
  const size_t nb = ceil_div(n, CP_TILE);
  DBuf cb((nb + 1) * sizeof(unsigned long long));
  unsigned long long* counts = cb.as<unsigned long long>();
  compact_count_kernel<<<unsigned(nb), CP_NT, 0, st>>>(pred, n, counts);
  FHIP_CHECK(hipGetLastError());
  scan_counts_kernel<<<1, 1024, 0, st>>>(counts, nb, counts + nb);
  FHIP_CHECK(hipGetLastError());
  unsigned long long total = 0;
  FHIP_CHECK(hipMemcpyAsync(&total, counts + nb, sizeof(total), hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  if (size_t(total) <= capacity && total) {
    compact_apply_kernel<<<unsigned(nb), CP_NT, 0, st>>>(pred, writer, n, counts);
    FHIP_CHECK(hipGetLastError());
  }
  return size_t(total);
}

inline int grid_for_n(size_t n, int per_cu = 8) {
  return int(std::max<size_t>(1, std::min<size_t>(ceil_div(n, 256), size_t(fill_blocks(256, per_cu)))));
}

// dst[i] = src[i * stride] for i < count (a strided read gathered into one
// contiguous block, so that a single linear copy brings it to the host).
__global__ __launch_bounds__(256) inline void gather_strided_u64_kernel(
    const unsigned long long* __restrict__ src, size_t stride, size_t count,
    unsigned long long* __restrict__ dst) {
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < count; i += size_t(gridDim.x) * 256)
    dst[i] = src[i * stride];
}

// Copies `count` 64-bit words spaced `stride` words apart from device memory
// into host memory `dst` and waits for the copy.
inline void copy_strided_to_host(const void* src, size_t stride, size_t count, void* dst,
                                 hipStream_t st) {
  if (count == 0) return;
  DBuf tmp(count * 8);
  gather_strided_u64_kernel<<<grid_for_n(count), 256, 0, st>>>(
      static_cast<const unsigned long long*>(src), stride, count, tmp.as<unsigned long long>());
  FHIP_CHECK(hipGetLastError());
  FHIP_CHECK(hipMemcpyAsync(dst, tmp.get(), count * 8, hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
}

template <class F>
void dispatch_type(int t, F&& f) {
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

}  // namespace fhip
