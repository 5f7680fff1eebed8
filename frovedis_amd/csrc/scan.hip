// prefix_sum (core/prefix_sum.hpp:67-88), set_separate / set_is_unique
// (core/set_operations.hpp:146-197) and batched lower/upper bound
// (core/lower_bound.hpp:13-101, core/upper_bound.hpp) for gfx950.
#include "fhip_device.hpp"

namespace fhip {
namespace {

constexpr int PS_NT = 256;
constexpr int PS_ITEMS = 16;
constexpr int PS_TILE = PS_NT * PS_ITEMS;

template <class T>
__global__ __launch_bounds__(PS_NT) void ps_reduce(const T* __restrict__ in, size_t n,
                                                   T* __restrict__ partial) {
  const size_t base = size_t(blockIdx.x) * PS_TILE;
  T s = 0;
#pragma unroll
  for (int j = 0; j < PS_ITEMS; j++) {
    size_t i = base + size_t(j) * PS_NT + threadIdx.x;
    if (i < n) s += in[i];
  }
  __shared__ T tmp[PS_NT / 64];
  T all;
  block_excl_scan<PS_NT>(s, tmp, &all);
  if (threadIdx.x == 0) partial[blockIdx.x] = all;
}

// Inclusive scan of one tile, offset by the scanned tile partials.
template <class T>
__global__ __launch_bounds__(PS_NT) void ps_tile(const T* __restrict__ in, T* __restrict__ out,
                                                 size_t n, const T* __restrict__ offs) {
  __shared__ T buf[PS_TILE];
  __shared__ T tmp[PS_NT / 64];
  const size_t base = size_t(blockIdx.x) * PS_TILE;
#pragma unroll
  for (int j = 0; j < PS_ITEMS; j++) {
    size_t i = base + size_t(j) * PS_NT + threadIdx.x;
    buf[j * PS_NT + threadIdx.x] = i < n ? in[i] : T(0);
  }
  __syncthreads();
  T* mine = buf + threadIdx.x * PS_ITEMS;
  T s = 0;
#pragma unroll
  for (int j = 0; j < PS_ITEMS; j++) s += mine[j];
  T e = block_excl_scan<PS_NT>(s, tmp, (T*)nullptr);
  T run = offs[blockIdx.x] + e;
#pragma unroll
  for (int j = 0; j < PS_ITEMS; j++) {
    run += mine[j];
    mine[j] = run;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < PS_ITEMS; j++) {
    size_t i = base + size_t(j) * PS_NT + threadIdx.x;
    if (i < n) out[i] = buf[j * PS_NT + threadIdx.x];
  }
}

template <class T>
void prefix_sum_t(const void* in, void* out, size_t n, hipStream_t st) {
  if (n == 0) return;
  const size_t nb = ceil_div(n, PS_TILE);
  DBuf pb((nb + 1) * sizeof(T));
  T* partial = pb.as<T>();
  ps_reduce<T><<<unsigned(nb), PS_NT, 0, st>>>((const T*)in, n, partial);
  FHIP_CHECK(hipGetLastError());
  scan_counts_kernel<T><<<1, 1024, 0, st>>>(partial, nb, partial + nb);
  FHIP_CHECK(hipGetLastError());
  ps_tile<T><<<unsigned(nb), PS_NT, 0, st>>>((const T*)in, (T*)out, n, partial);
  FHIP_CHECK(hipGetLastError());
}

template <class T>
struct RunStart {
  const T* k;
  __device__ bool operator()(size_t i) const { return i == 0 || k[i] != k[i - 1]; }
};

template <class T>
__global__ void sep_finish(size_t* out, size_t count, size_t n) {
  out[count] = n;
}

template <class T>
size_t set_separate_t(const void* sorted, size_t n, size_t* out, hipStream_t st) {
  if (n == 0) {
    size_t z = 0;
    FHIP_CHECK(hipMemcpyAsync(out, &z, sizeof(z), hipMemcpyHostToDevice, st));
    FHIP_CHECK(hipStreamSynchronize(st));
    return 1;
  }
  size_t c = compact_indices(RunStart<T>{(const T*)sorted}, n, out, st);
  sep_finish<T><<<1, 1, 0, st>>>(out, c, n);
  FHIP_CHECK(hipGetLastError());
  return c + 1;
}

template <class T>
__global__ void dup_flag_kernel(const T* __restrict__ k, size_t n, int* flag) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x + 1; i < n; i += stride)
    if (!(k[i] != k[i - 1])) *flag = 0;
}

template <class T, bool UPPER>
__global__ void bound_kernel(const T* __restrict__ s, size_t m, const T* __restrict__ q,
                             size_t nq, size_t* __restrict__ out) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nq; i += stride) {
    const T x = q[i];
    size_t lo = 0, hi = m;
    // the reference's comparisons (core/lower_bound.hpp:44-48,
    // core/upper_bound.hpp:44-48): lower goes left when x <= s[mid], upper
    // goes right when x >= s[mid] (a NaN query: m and 0)
    while (lo < hi) {
      size_t mid = (lo + hi) >> 1;
      bool go_right = UPPER ? (x >= s[mid]) : !(x <= s[mid]);
      if (go_right) lo = mid + 1; else hi = mid;
    }
    out[i] = lo;
  }
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

}  // namespace

void prefix_sum(int type, const void* in, void* out, size_t n, hipStream_t st) {
  dispatch(type, [&](auto z) { prefix_sum_t<decltype(z)>(in, out, n, st); });
}

size_t set_separate(int type, const void* sorted, size_t n, size_t* out, hipStream_t st) {
  size_t r = 0;
  dispatch(type, [&](auto z) { r = set_separate_t<decltype(z)>(sorted, n, out, st); });
  return r;
}

}  // namespace fhip

extern "C" {

int fhip_dev_prefix_sum(int type, const void* in, void* out, size_t n, void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(n == 0 || (in && out), FHIP_ERR_ARG, "prefix_sum: null pointer");
    fhip::prefix_sum(type, in, out, n, fhip::as_stream(stream));
  });
}

int fhip_prefix_sum(int type, const void* in, void* out, size_t n) {
  return fhip::guarded([&] {
    if (n == 0) return;
    const size_t b = size_t(fhip::type_bytes(type)) * n;
    fhip::DBuf di(b), dout(b);
    hipStream_t st = fhip::as_stream(nullptr);
    fhip::copy_h2d(di.get(), in, b, st);
    fhip::prefix_sum(type, di.get(), dout.get(), n, st);
    fhip::copy_d2h(out, dout.get(), b, st);
  });
}

int fhip_dev_set_separate(int type, const void* sorted, size_t n, size_t* out,
                          size_t* out_n, void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(out && out_n && (n == 0 || sorted), FHIP_ERR_ARG, "set_separate: null pointer");
    *out_n = fhip::set_separate(type, sorted, n, out, fhip::as_stream(stream));
  });
}

int fhip_set_separate(int type, const void* sorted, size_t n, size_t* out, size_t* out_n) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(out && out_n, FHIP_ERR_ARG, "set_separate: null pointer");
    const size_t b = size_t(fhip::type_bytes(type)) * n;
    fhip::DBuf dk(b), dout((n + 1) * sizeof(size_t));
    hipStream_t st = fhip::as_stream(nullptr);
    fhip::copy_h2d(dk.get(), sorted, b, st);
    size_t c = fhip::set_separate(type, dk.get(), n, dout.as<size_t>(), st);
    fhip::copy_d2h(out, dout.get(), c * sizeof(size_t), st);
    *out_n = c;
  });
}

int fhip_set_is_unique(int type, const void* sorted, size_t n, int* is_unique) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(is_unique && (n == 0 || sorted), FHIP_ERR_ARG, "set_is_unique: null pointer");
    const size_t b = size_t(fhip::type_bytes(type)) * n;
    fhip::DBuf dk(b);
    hipStream_t st = fhip::as_stream(nullptr);
    fhip::copy_h2d(dk.get(), sorted, b, st);
    const int rc = fhip_dev_set_is_unique(type, dk.get(), n, is_unique, nullptr);
    if (rc != FHIP_OK) throw fhip::Error(rc, fhip_last_error());
  });
}

int fhip_dev_set_is_unique(int type, const void* sorted, size_t n, int* is_unique,
                           void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(is_unique, FHIP_ERR_ARG, "set_is_unique: null pointer");
    hipStream_t st = fhip::as_stream(stream);
    *is_unique = 1;
    if (n < 2) return;
    fhip::DBuf fb(sizeof(int));
    int one = 1;
    FHIP_CHECK(hipMemcpyAsync(fb.get(), &one, sizeof(int), hipMemcpyHostToDevice, st));
    int grid = int(std::min<size_t>(fhip::ceil_div(n, 256), size_t(fhip::fill_blocks(256, 4))));
    fhip::dispatch(type, [&](auto z) {
      using T = decltype(z);
      fhip::dup_flag_kernel<T><<<grid, 256, 0, st>>>((const T*)sorted, n, fb.as<int>());
    });
    FHIP_CHECK(hipGetLastError());
    FHIP_CHECK(hipMemcpyAsync(is_unique, fb.get(), sizeof(int), hipMemcpyDeviceToHost, st));
    FHIP_CHECK(hipStreamSynchronize(st));
  });
}

int fhip_dev_lower_bound(int type, const void* sorted, size_t m, const void* queries,
                         size_t nq, size_t* out, int upper, void* stream) {
  return fhip::guarded([&] {
    if (nq == 0) return;
    FHIP_REQUIRE(queries && out && (m == 0 || sorted), FHIP_ERR_ARG, "lower_bound: null pointer");
    hipStream_t st = fhip::as_stream(stream);
    int grid = int(std::min<size_t>(fhip::ceil_div(nq, 256), size_t(fhip::fill_blocks(256, 8))));
    fhip::dispatch(type, [&](auto z) {
      using T = decltype(z);
      if (upper)
        fhip::bound_kernel<T, true><<<grid, 256, 0, st>>>((const T*)sorted, m, (const T*)queries, nq, out);
      else
        fhip::bound_kernel<T, false><<<grid, 256, 0, st>>>((const T*)sorted, m, (const T*)queries, nq, out);
    });
    FHIP_CHECK(hipGetLastError());
  });
}

}  // extern "C"
