// equi_join on one rank's share for gfx950 (dataframe/join.hpp:304-326 VE
// body, join.cc:10-61 shrink_missed, multi_equi_join :198-302).
//
// strategy FHIP_JOIN_AUTO follows the reference: a hash table (for integer
// keys the joins' own open-addressing table below, otherwise a
// unique_hashtable) is built on RIGHT with the right row ids as values
// (skipped when an evenly spaced sample of RIGHT already repeats a key,
// which proves it not unique); if the right keys are unique, LEFT is probed
// in order and the hits are compacted (hash join, output in left order);
// otherwise both sides are stable-sorted and every left row (in sorted
// order) is paired with its run of equal right keys (sort-merge, output in
// (key, left position, right position) order).
// FHIP_JOIN_HASH requires unique right keys; FHIP_JOIN_SORTMERGE skips the
// table build.
#include "densejoin.hpp"
#include "fhip_device.hpp"
#include "hashtable.hpp"
#include "radix_sort.hpp"

#include <vector>

namespace fhip {

void prefix_sum(int type, const void* in, void* out, size_t n, hipStream_t st);

namespace {

struct HitPred {
  const uint8_t* f;
  __device__ bool operator()(size_t i) const { return f[i] != 0; }
};

struct HitWriter {
  const size_t* lidx;
  const uint64_t* vals;
  size_t* lo;
  size_t* ro;
  __device__ void operator()(size_t pos, size_t i) const {
    lo[pos] = lidx[i];
    ro[pos] = vals[i];
  }
};

struct MissPred {
  const uint8_t* f;
  __device__ bool operator()(size_t i) const { return f[i] == 0; }
};

struct NoMatchPred {
  const unsigned long long* cnt;
  __device__ bool operator()(size_t i) const { return cnt[i] == 0; }
};

// i -> the row id at position i
struct RowIdMap {
  const size_t* ids;
  __device__ size_t operator()(size_t i) const { return ids[i]; }
};

// number of right matches of every sorted left key, and where they start.
// One workgroup per block of MC_NT sorted left keys.  The blocks' right
// ranges (a lower bound of the block's first key and an upper bound of its
// last, over all of RIGHT) come first, one thread per block, so those long
// searches run side by side instead of holding each block; a block whose
// range has at most MC_CAP keys stages it in LDS and every key searches
// there, a wider one searches its range in memory.
constexpr int MC_NT = 256;
constexpr int MC_CAP = 4096;

template <class K>
__device__ __forceinline__ size_t lower_in(const K* r, size_t lo, size_t hi, K x) {
  while (lo < hi) {
    const size_t mid = (lo + hi) >> 1;
    if (r[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

template <class K>
__device__ __forceinline__ size_t upper_in(const K* r, size_t lo, size_t hi, K x) {
  while (lo < hi) {
    const size_t mid = (lo + hi) >> 1;
    if (x < r[mid]) hi = mid; else lo = mid + 1;
  }
  return lo;
}

template <class K>
__global__ __launch_bounds__(256) void match_range_kernel(const K* __restrict__ lk, size_t nl,
                                                          const K* __restrict__ rk, size_t nr,
                                                          size_t nblocks,
                                                          unsigned long long* __restrict__ rng) {
  const size_t b = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (b >= nblocks) return;
  const size_t b0 = b * MC_NT;
  const size_t b1 = b0 + MC_NT < nl ? b0 + MC_NT : nl;
  rng[2 * b] = lower_in(rk, 0, nr, lk[b0]);
  rng[2 * b + 1] = upper_in(rk, 0, nr, lk[b1 - 1]);
}

template <class K>
__global__ __launch_bounds__(MC_NT) void match_count_kernel(const K* __restrict__ lk, size_t nl,
                                                            const K* __restrict__ rk, size_t nr,
                                                            const unsigned long long* __restrict__ rng,
                                                            unsigned long long* __restrict__ cnt,
                                                            unsigned long long* __restrict__ first) {
  __shared__ K sr[MC_CAP];
  const size_t b0 = size_t(blockIdx.x) * MC_NT;
  const size_t b1 = b0 + MC_NT < nl ? b0 + MC_NT : nl;
  size_t r0 = 0, r1 = nr;
  // floating keys (NaN, -0.0) are not totally ordered by <: search all
  if constexpr (std::is_integral<K>::value) {
    r0 = size_t(rng[2 * size_t(blockIdx.x)]);
    r1 = size_t(rng[2 * size_t(blockIdx.x) + 1]);
  }
  const bool staged = std::is_integral<K>::value && r1 - r0 <= size_t(MC_CAP);
  if (staged)
    for (size_t j = threadIdx.x; j < r1 - r0; j += MC_NT) sr[j] = rk[r0 + j];
  __syncthreads();
  const size_t i = b0 + threadIdx.x;
  if (i >= b1) return;
  const K x = lk[i];
  size_t a, e;
  if (staged) {
    a = lower_in(sr, size_t(0), r1 - r0, x);
    e = upper_in(sr, a, r1 - r0, x);
    a += r0;
    e += r0;
  } else {
    a = lower_in(rk, r0, r1, x);
    e = upper_in(rk, a, r1, x);
  }
  cnt[i] = e - a;
  first[i] = a;
}

// Writes the pairs of one block of PW_NT sorted left rows: the block's
// output is the contiguous range [incl[b0] - cnt[b0], incl[b1 - 1]); lanes
// take consecutive output positions (coalesced stores) and find their left
// row by a binary search over the block's inclusive counts in LDS.
constexpr int PW_NT = 256;

__global__ __launch_bounds__(PW_NT) void pair_write_kernel(
    const unsigned long long* __restrict__ incl, const unsigned long long* __restrict__ cnt,
    const unsigned long long* __restrict__ first, const size_t* __restrict__ lidx_sorted,
    const size_t* __restrict__ ridx_sorted, size_t nl, size_t* __restrict__ lo,
    size_t* __restrict__ ro) {
  __shared__ unsigned long long s_incl[PW_NT];
  __shared__ unsigned long long s_first[PW_NT];
  __shared__ size_t s_l[PW_NT];
  const size_t b0 = size_t(blockIdx.x) * PW_NT;
  const unsigned m = unsigned(b0 + PW_NT < nl ? PW_NT : nl - b0);
  const unsigned t = threadIdx.x;
  if (t < m) {
    s_incl[t] = incl[b0 + t];
    s_first[t] = first[b0 + t];
    s_l[t] = lidx_sorted[b0 + t];
  }
  __syncthreads();
  const unsigned long long out0 = s_incl[0] - cnt[b0];
  const unsigned long long out1 = s_incl[m - 1];
  for (unsigned long long p = out0 + t; p < out1; p += PW_NT) {
    // first row j of the block with s_incl[j] > p
    unsigned a = 0, z = m;
    while (a < z) {
      const unsigned mid = (a + z) >> 1;
      if (s_incl[mid] > p) z = mid; else a = mid + 1;
    }
    const unsigned long long row_start = a ? s_incl[a - 1] : out0;
    lo[p] = s_l[a];
    ro[p] = ridx_sorted[s_first[a] + (p - row_start)];
  }
}

// The pairs of the narrow sort-merge path: as pair_write_kernel, with the
// sorted sides given as rows (0..n-1) of consecutive row ids ids[0] + row.
__global__ __launch_bounds__(PW_NT) void pair_write_rows_kernel(
    const unsigned long long* __restrict__ incl, const unsigned long long* __restrict__ cnt,
    const unsigned long long* __restrict__ first, const uint32_t* __restrict__ lrow,
    const uint32_t* __restrict__ rrow, const size_t* __restrict__ lidx,
    const size_t* __restrict__ ridx, size_t nl, size_t* __restrict__ lo, size_t* __restrict__ ro) {
  __shared__ unsigned long long s_incl[PW_NT];
  __shared__ unsigned long long s_first[PW_NT];
  __shared__ size_t s_l[PW_NT];
  const size_t b0 = size_t(blockIdx.x) * PW_NT;
  const unsigned m = unsigned(b0 + PW_NT < nl ? PW_NT : nl - b0);
  const unsigned t = threadIdx.x;
  const size_t l0 = lidx[0], r0 = ridx[0];
  if (t < m) {
    s_incl[t] = incl[b0 + t];
    s_first[t] = first[b0 + t];
    s_l[t] = l0 + lrow[b0 + t];
  }
  __syncthreads();
  const unsigned long long out0 = s_incl[0] - cnt[b0];
  const unsigned long long out1 = s_incl[m - 1];
  for (unsigned long long p = out0 + t; p < out1; p += PW_NT) {
    unsigned a = 0, z = m;
    while (a < z) {
      const unsigned mid = (a + z) >> 1;
      if (s_incl[mid] > p) z = mid; else a = mid + 1;
    }
    const unsigned long long row_start = a ? s_incl[a - 1] : out0;
    lo[p] = s_l[a];
    ro[p] = r0 + rrow[s_first[a] + (p - row_start)];
  }
}

// i -> ids[0] + rows[i] (a sorted side of the narrow path as row ids)
struct RowOffMap {
  const size_t* ids;
  const uint32_t* rows;
  __device__ size_t operator()(size_t i) const { return ids[0] + rows[i]; }
};

// an evenly spaced sample of the right keys
template <class K>
__global__ void sample_kernel(const K* __restrict__ keys, size_t step, size_t m,
                              K* __restrict__ out) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += stride)
    out[i] = keys[i * step];
}

template <class K>
__global__ void sample_dup_kernel(const K* __restrict__ k, size_t m, int* __restrict__ flag) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x + 1; i < m; i += stride)
    if (k[i] == k[i - 1]) *flag = 1;
}

// True when an evenly spaced sample of the right keys already holds two
// equal keys: then the right side is certainly not unique and the table
// build the reference uses to find that out can be skipped.  No duplicate
// in the sample decides nothing (the table build decides).
template <class K>
bool sample_has_duplicate(int ktype, const K* right, size_t nr, hipStream_t st) {
  constexpr size_t kSample = 1 << 16;
  if (nr < 2 * kSample) return false;
  const size_t step = nr / kSample;
  DBuf smp(kSample * sizeof(K)), flag(sizeof(int));
  sample_kernel<K><<<grid_for_n(kSample), 256, 0, st>>>(right, step, kSample, smp.as<K>());
  FHIP_CHECK(hipGetLastError());
  sort_pairs(ktype, smp.get(), smp.get(), 0, nullptr, nullptr, kSample, false, false, st);
  FHIP_CHECK(hipMemsetAsync(flag.get(), 0, sizeof(int), st));
  sample_dup_kernel<K><<<grid_for_n(kSample), 256, 0, st>>>(smp.as<K>(), kSample, flag.as<int>());
  FHIP_CHECK(hipGetLastError());
  int f = 0;
  FHIP_CHECK(hipMemcpyAsync(&f, flag.get(), sizeof(int), hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  return f != 0;
}

// only_left (may be null): receives the left row ids without a match in
// stable key order of LEFT (outer_equi_join, join.hpp:352-378), *n_only
// their count.
// inplace: the inputs themselves are sorted (with their indices), as the
// VE body leaves them (join.hpp:320-322); else sorted copies are used.
// The sort-merge join when both sides' row ids are consecutive and each
// side's keys share their bits above the low 32: both sides sorted as (low
// 32 key bits, 32-bit row) pairs (sort_narrow: half the bytes per pass of
// the key/row-id sort), merged on the 32-bit keys, and the pairs written as
// ids[0] + row.  Same pairs in the same (key, left, right) order as the
// general path.  Returns false, having written nothing the caller keeps,
// when a side does not qualify.
template <class K>
bool sort_merge_narrow(int ktype, const K* left, const size_t* lidx, size_t nl, const K* right,
                       const size_t* ridx, size_t nr, size_t* lo, size_t* ro, size_t capacity,
                       size_t* only_left, size_t* n_only, hipStream_t st, size_t* total_out) {
  DBuf lk(nl * 4), lr(nl * 4), rk(nr * 4), rr(nr * 4), bad(sizeof(unsigned));
  FHIP_CHECK(hipMemsetAsync(bad.get(), 0, sizeof(unsigned), st));
  unsigned hl = 0, hr = 0;
  // radix_sort(right, right_idx); radix_sort(left, left_idx) (join.hpp:320-322)
  if (!sort_narrow(ktype, right, ridx, nr, rk.as<uint32_t>(), rr.as<uint32_t>(), &hr,
                   bad.as<unsigned>(), st))
    return false;
  if (!sort_narrow(ktype, left, lidx, nl, lk.as<uint32_t>(), lr.as<uint32_t>(), &hl,
                   bad.as<unsigned>(), st))
    return false;
  DBuf cnt(nl * sizeof(unsigned long long)), first(nl * sizeof(unsigned long long)),
      incl(nl * sizeof(unsigned long long));
  unsigned long long* rb = static_cast<unsigned long long*>(pinned_scratch(0, 16));
  rb[0] = 0;
  if (hl == hr) {
    {
      ProfScope prof("join_merge", st);
      const size_t nb = ceil_div(nl, size_t(MC_NT));
      DBuf rng(2 * nb * 8);
      match_range_kernel<uint32_t><<<unsigned(ceil_div(nb, size_t(256))), 256, 0, st>>>(
          lk.as<uint32_t>(), nl, rk.as<uint32_t>(), nr, nb, rng.as<unsigned long long>());
      FHIP_CHECK(hipGetLastError());
      match_count_kernel<uint32_t><<<unsigned(nb), MC_NT, 0, st>>>(
          lk.as<uint32_t>(), nl, rk.as<uint32_t>(), nr, rng.as<unsigned long long>(),
          cnt.as<unsigned long long>(), first.as<unsigned long long>());
      FHIP_CHECK(hipGetLastError());
    }
    prefix_sum(FHIP_U64, cnt.get(), incl.get(), nl, st);
    FHIP_CHECK(hipMemcpyAsync(rb, incl.as<unsigned long long>() + (nl - 1), 8,
                              hipMemcpyDeviceToHost, st));
  } else {
    // the sides differ above the low 32 bits: no key matches
    FHIP_CHECK(hipMemsetAsync(cnt.get(), 0, nl * sizeof(unsigned long long), st));
  }
  FHIP_CHECK(hipMemcpyAsync(rb + 1, bad.get(), sizeof(unsigned), hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  if ((rb[1] & 0xFFFFFFFFull) != 0) return false;  // row ids not consecutive
  const unsigned long long total = rb[0];
  if (total && size_t(total) <= capacity) {
    ProfScope prof("join_write", st);
    pair_write_rows_kernel<<<unsigned(ceil_div(nl, PW_NT)), PW_NT, 0, st>>>(
        incl.as<unsigned long long>(), cnt.as<unsigned long long>(),
        first.as<unsigned long long>(), lr.as<uint32_t>(), rr.as<uint32_t>(), lidx, ridx, nl,
        lo, ro);
    FHIP_CHECK(hipGetLastError());
  }
  if (only_left)
    *n_only = compact(NoMatchPred{cnt.as<unsigned long long>()},
                      RowOffMap{lidx, lr.as<uint32_t>()}, nl, only_left, st);
  *total_out = size_t(total);
  return true;
}

template <class K>
size_t sort_merge_t(int ktype, const K* left, const size_t* lidx, size_t nl, const K* right,
                    const size_t* ridx, size_t nr, size_t* lo, size_t* ro, size_t capacity,
                    size_t* only_left, size_t* n_only, hipStream_t st, bool inplace = false) {
  if (n_only) *n_only = 0;
  if constexpr (std::is_integral<K>::value) {
    // the narrow path (both row-id columns consecutive, each side's keys
    // within one 2^32 stretch); the general one below otherwise
    size_t total = 0;
    if (!inplace && nl > 0 && nr > 0 && nl < (size_t(1) << 32) && nr < (size_t(1) << 32) &&
        tune_get("join.narrow", 1) != 0 &&
        sort_merge_narrow<K>(ktype, left, lidx, nl, right, ridx, nr, lo, ro, capacity, only_left,
                             n_only, st, &total))
      return total;
    if (n_only) *n_only = 0;
  }
  // radix_sort(right, right_idx); radix_sort(left, left_idx) (join.hpp:320-322)
  if (inplace) {
    if (nr) sort_pairs(ktype, right, const_cast<K*>(right), 8, ridx, const_cast<size_t*>(ridx),
                       nr, false, false, st);
    if (nl) sort_pairs(ktype, left, const_cast<K*>(left), 8, lidx, const_cast<size_t*>(lidx), nl,
                       false, false, st);
  }
  if (nl == 0) return 0;
  if (nr == 0 && !only_left) return 0;
  DBuf lk, li, rk, ri;
  if (!inplace) {
    lk.reset(nl * sizeof(K));
    li.reset(nl * sizeof(size_t));
    rk.reset(nr * sizeof(K));
    ri.reset(nr * sizeof(size_t));
    if (nr) sort_pairs(ktype, right, rk.get(), 8, ridx, ri.get(), nr, false, false, st);
    sort_pairs(ktype, left, lk.get(), 8, lidx, li.get(), nl, false, false, st);
  }
  const K* lks = inplace ? left : lk.as<K>();
  const size_t* lis = inplace ? lidx : li.as<size_t>();
  const K* rks = inplace ? right : rk.as<K>();
  const size_t* ris = inplace ? ridx : ri.as<size_t>();
  if (nr == 0) {
    FHIP_CHECK(hipMemcpyAsync(only_left, lis, nl * sizeof(size_t), hipMemcpyDeviceToDevice,
                              st));
    FHIP_CHECK(hipStreamSynchronize(st));
    *n_only = nl;
    return 0;
  }
  DBuf cnt(nl * sizeof(unsigned long long)), first(nl * sizeof(unsigned long long)),
      incl(nl * sizeof(unsigned long long));
  {
    ProfScope prof("join_merge", st);
    const size_t nb = ceil_div(nl, size_t(MC_NT));
    DBuf rng(2 * nb * 8);
    if (std::is_integral<K>::value) {
      match_range_kernel<K><<<unsigned(ceil_div(nb, size_t(256))), 256, 0, st>>>(
          lks, nl, rks, nr, nb, rng.as<unsigned long long>());
      FHIP_CHECK(hipGetLastError());
    }
    match_count_kernel<K><<<unsigned(nb), MC_NT, 0, st>>>(
        lks, nl, rks, nr, rng.as<unsigned long long>(), cnt.as<unsigned long long>(),
        first.as<unsigned long long>());
    FHIP_CHECK(hipGetLastError());
  }
  prefix_sum(FHIP_U64, cnt.get(), incl.get(), nl, st);
  unsigned long long total = 0;
  FHIP_CHECK(hipMemcpyAsync(&total, incl.as<unsigned long long>() + (nl - 1), sizeof(total),
                            hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  if (total && size_t(total) <= capacity) {
    ProfScope prof("join_write", st);
    pair_write_kernel<<<unsigned(ceil_div(nl, PW_NT)), PW_NT, 0, st>>>(
        incl.as<unsigned long long>(), cnt.as<unsigned long long>(),
        first.as<unsigned long long>(), lis, ris, nl, lo, ro);
    FHIP_CHECK(hipGetLastError());
  }
  if (only_left)
    *n_only = compact(NoMatchPred{cnt.as<unsigned long long>()}, RowIdMap{lis}, nl,
                      only_left, st);
  return size_t(total);
}

// ---- the joins' own table for integer keys -------------------------------------
// A join only looks keys up, so for integer keys it does not need the
// unique_hashtable layout: open addressing in HBM, 16-byte {key, value}
// slots, at most 50 % full (the tunable "join.load_pct", 10 to 50: with
// 2^27 unique keys, 50 % halves the 8 GB table of 25 % and its clearing, and
// the join of 2^29 probes took 30.2 ms against 31.1-31.8), home slot by a
// multiply-shift hash, linear
// probing, insertion by a 64-bit atomicCAS on the key word (which also finds
// a repeated key: then the right side is not unique, as the reference's
// is_unique_ok reports).  The all-ones key, the empty marker, is kept aside
// in a side record.  Floating keys keep the reference layout, whose bit
// pattern versus == behaviour (-0.0, NaN) the result depends on.
constexpr unsigned long long JT_EMPTY = ~0ull;
constexpr int JT_NT = 256;
constexpr int JT_PROBE_U = 4;  // queries per thread and step of the probe

template <class K>
__device__ __forceinline__ unsigned long long jt_bits(K k) {
  return static_cast<unsigned long long>(static_cast<typename std::make_unsigned<K>::type>(k));
}

__device__ __forceinline__ unsigned long long jt_home(unsigned long long k, int shift) {
  return (k * 0x9E3779B97F4A7C15ull) >> shift;
}

// side = {rows with the empty-marker key, the value of the first, repeated key}
// The rest of an insertion whose home slot h held another key.
__device__ __noinline__ void jt_insert_rest(unsigned long long* slots, unsigned long long mask,
                                            unsigned long long k, unsigned long long v,
                                            unsigned long long h, unsigned long long* side) {
  while (true) {
    h = (h + 1) & mask;
    const unsigned long long prev = atomicCAS(&slots[2 * h], JT_EMPTY, k);
    if (prev == JT_EMPTY) {
      slots[2 * h + 1] = v;
      return;
    }
    if (prev == k) {
      atomicOr(&side[2], 1ull);
      return;
    }
  }
}

template <class K>
__global__ __launch_bounds__(JT_NT) void jt_insert_kernel(const K* __restrict__ keys,
                                                          const unsigned long long* __restrict__ vals,
                                                          size_t n, unsigned long long* slots,
                                                          unsigned long long mask, int shift,
                                                          unsigned long long* side) {
  // JT_PROBE_U keys per thread and step, their first CAS issued together
  constexpr int U = JT_PROBE_U;
  const size_t stride = size_t(gridDim.x) * JT_NT * U;
  for (size_t i0 = size_t(blockIdx.x) * JT_NT * U + threadIdx.x; i0 < n; i0 += stride) {
    unsigned long long k[U], v[U], h[U], prev[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const size_t i = i0 + size_t(u) * JT_NT;
      k[u] = i < n ? jt_bits(keys[i]) : JT_EMPTY;
      v[u] = i < n ? vals[i] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      h[u] = jt_home(k[u], shift);
      prev[u] = k[u] != JT_EMPTY ? atomicCAS(&slots[2 * h[u]], JT_EMPTY, k[u]) : JT_EMPTY;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const size_t i = i0 + size_t(u) * JT_NT;
      if (i >= n) continue;
      if (k[u] == JT_EMPTY) {
        if (atomicAdd(&side[0], 1ull) == 0) side[1] = v[u];
        else atomicOr(&side[2], 1ull);
      } else if (prev[u] == JT_EMPTY) {
        slots[2 * h[u] + 1] = v[u];
      } else if (prev[u] == k[u]) {
        atomicOr(&side[2], 1ull);
      } else {  // a collision: walk on from the next slot
        jt_insert_rest(slots, mask, k[u], v[u], h[u], side);
      }
    }
  }
}

// The rest of a linear probe that started at slot h and did not end there:
// {1, value} when k is found, {0, 0} at the first empty slot.
__device__ __noinline__ ulonglong2 jt_probe_rest(const unsigned long long* __restrict__ slots,
                                                 unsigned long long mask, unsigned long long k,
                                                 unsigned long long h) {
  while (true) {
    h = (h + 1) & mask;
    const ulonglong2 e = *reinterpret_cast<const ulonglong2*>(slots + 2 * h);
    if (e.x == k) return make_ulonglong2(1ull, e.y);
    if (e.x == JT_EMPTY) return make_ulonglong2(0ull, 0ull);
  }
}

template <class K>
__global__ __launch_bounds__(JT_NT) void jt_probe_kernel(const K* __restrict__ q, size_t m,
                                                         const unsigned long long* __restrict__ slots,
                                                         unsigned long long mask, int shift,
                                                         const unsigned long long* __restrict__ side,
                                                         unsigned long long* __restrict__ vals_out,
                                                         uint8_t* __restrict__ found) {
  // JT_PROBE_U queries per thread and step, their first slot reads issued
  // together (each lookup is a dependent read chain: query, then slot)
  constexpr int U = JT_PROBE_U;
  const size_t stride = size_t(gridDim.x) * JT_NT * U;
  for (size_t j0 = size_t(blockIdx.x) * JT_NT * U + threadIdx.x; j0 < m; j0 += stride) {
    unsigned long long k[U], h[U];
    ulonglong2 e[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const size_t j = j0 + size_t(u) * JT_NT;
      k[u] = j < m ? jt_bits(q[j]) : JT_EMPTY;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      h[u] = jt_home(k[u], shift);
      e[u] = k[u] != JT_EMPTY ? *reinterpret_cast<const ulonglong2*>(slots + 2 * h[u])
                              : make_ulonglong2(JT_EMPTY, 0ull);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const size_t j = j0 + size_t(u) * JT_NT;
      bool f = false;
      unsigned long long v = 0;
      if (k[u] == JT_EMPTY) {
        f = side[0] != 0;
        v = f ? side[1] : 0ull;
      } else if (e[u].x == k[u]) {
        f = true;
        v = e[u].y;
      } else if (e[u].x != JT_EMPTY) {  // a collision: walk on from the next slot
        const ulonglong2 r = jt_probe_rest(slots, mask, k[u], h[u]);
        f = r.x != 0;
        v = r.y;
      }
      if (j < m) {
        vals_out[j] = v;
        found[j] = f ? 1 : 0;
      }
    }
  }
}

// The probe fused with the compaction of its hits (inner joins): a
// workgroup takes the next tile of JP_NT * JP_ITEMS queries in arrival
// order, looks every query up (the ITEMS slot reads of a thread in flight
// together), counts its hits, finds the hits of all earlier tiles by
// decoupled look-back and writes its pairs (left row id, looked-up value)
// at once, in query order.  No per-query value or flag array is written and
// none is read back.  Pairs at positions >= capacity are counted, not
// written; tile_state (tiles words), ticket and total start at zero.
constexpr int JP_NT = 1024;
constexpr int JP_ITEMS = 8;

template <class K>
__global__ __launch_bounds__(JP_NT) void jt_probe_compact_kernel(
    const K* __restrict__ q, size_t m, const unsigned long long* __restrict__ slots,
    unsigned long long mask, int shift, const unsigned long long* __restrict__ side,
    const size_t* __restrict__ lidx, size_t* __restrict__ lo, size_t* __restrict__ ro,
    size_t capacity, unsigned long long* tile_state, unsigned* ticket,
    unsigned long long* total) {
  constexpr int WAVES = JP_NT / 64;
  constexpr int COUNTS = JP_ITEMS * WAVES;  // (round, wave) counts, round-major
  constexpr int PER_LANE = COUNTS / 64;
  static_assert(COUNTS % 64 == 0, "the counts split evenly over one wave");
  __shared__ unsigned s_cnt[COUNTS];
  __shared__ unsigned s_tile;
  __shared__ unsigned long long s_excl;
  const unsigned lane = __lane_id();
  const unsigned wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const unsigned tile = s_tile;
  const size_t base = size_t(tile) * (JP_NT * JP_ITEMS);
  unsigned long long k[JP_ITEMS], v[JP_ITEMS];
  bool in[JP_ITEMS];
#pragma unroll
  for (int j = 0; j < JP_ITEMS; j++) {
    const size_t i = base + size_t(j) * JP_NT + threadIdx.x;
    in[j] = i < m;
    k[j] = in[j] ? jt_bits(q[i]) : 0ull;
  }
  // home slots of every query read together
  ulonglong2 e[JP_ITEMS];
#pragma unroll
  for (int j = 0; j < JP_ITEMS; j++)
    e[j] = (in[j] && k[j] != JT_EMPTY)
               ? *reinterpret_cast<const ulonglong2*>(slots + 2 * jt_home(k[j], shift))
               : make_ulonglong2(JT_EMPTY, 0ull);
  unsigned long long hit[JP_ITEMS];
#pragma unroll
  for (int j = 0; j < JP_ITEMS; j++) {
    bool f = false;
    v[j] = 0;
    if (in[j]) {
      if (k[j] == JT_EMPTY) {
        f = side[0] != 0;
        v[j] = f ? side[1] : 0ull;
      } else if (e[j].x == k[j]) {
        f = true;
        v[j] = e[j].y;
      } else if (e[j].x != JT_EMPTY) {  // a collision: walk on from the next slot
        const ulonglong2 r = jt_probe_rest(slots, mask, k[j], jt_home(k[j], shift));
        f = r.x != 0;
        v[j] = r.y;
      }
    }
    hit[j] = __ballot(f);
  }
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < JP_ITEMS; j++) s_cnt[j * WAVES + wave] = unsigned(__popcll(hit[j]));
  }
  __syncthreads();
  if (wave == 0) {
    unsigned c[PER_LANE];
    unsigned sum = 0;
#pragma unroll
    for (int t = 0; t < PER_LANE; t++) {
      c[t] = s_cnt[lane * PER_LANE + t];
      sum += c[t];
    }
    const unsigned incl = wave_incl_scan(sum);
    const unsigned run = __shfl(incl, 63, 64);
    unsigned ex = incl - sum;
#pragma unroll
    for (int t = 0; t < PER_LANE; t++) {
      s_cnt[lane * PER_LANE + t] = ex;
      ex += c[t];
    }
    const unsigned long long excl = compact_lookback(tile_state, tile, run);
    if (lane == 0) {
      if (tile == gridDim.x - 1) *total = excl + run;
      s_excl = excl;
    }
  }
  __syncthreads();
  const unsigned long long excl = s_excl;
#pragma unroll
  for (int j = 0; j < JP_ITEMS; j++) {
    if ((hit[j] >> lane) & 1ull) {
      const size_t pos = size_t(excl + s_cnt[j * WAVES + wave] + lane_rank(hit[j]));
      if (pos < capacity) {
        lo[pos] = lidx[base + size_t(j) * JP_NT + threadIdx.x];
        ro[pos] = v[j];
      }
    }
  }
}

// The joins' table over (keys, vals) for integer keys.
struct JoinTable {
  DBuf tab, side;
  unsigned long long nslots = 0;
  int bits = 0;
  bool unique = true;
};

// Builds the table; unique is false when the keys repeat (the table is
// then not usable).
template <class K>
void jt_build(JoinTable& t, const K* keys, const unsigned long long* vals, size_t n,
              hipStream_t st) {
  const long long load_pct = std::max(10LL, std::min(50LL, tune_get("join.load_pct", 50)));
  int bits = 10;
  while ((size_t(1) << bits) * size_t(load_pct) < 100 * n) bits++;
  t.bits = bits;
  t.nslots = 1ull << bits;
  t.tab.reset(size_t(t.nslots) * 16);
  t.side.reset(3 * sizeof(unsigned long long));
  FHIP_CHECK(hipMemsetAsync(t.tab.get(), 0xFF, size_t(t.nslots) * 16, st));
  FHIP_CHECK(hipMemsetAsync(t.side.get(), 0, 3 * sizeof(unsigned long long), st));
  {
    ProfScope prof("join_build", st);
    if (n)
      jt_insert_kernel<K><<<grid_for_n(n), JT_NT, 0, st>>>(
          keys, vals, n, t.tab.as<unsigned long long>(), t.nslots - 1, 64 - bits,
          t.side.as<unsigned long long>());
    FHIP_CHECK(hipGetLastError());
  }
  unsigned long long dup = 0;
  FHIP_CHECK(hipMemcpyAsync(&dup, t.side.as<unsigned long long>() + 2, sizeof(dup),
                            hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  t.unique = dup == 0;
}

// Probes q: vals_out / found per query.
template <class K>
void jt_probe(const JoinTable& t, const K* q, size_t m, unsigned long long* vals_out,
              uint8_t* found, hipStream_t st) {
  if (!m) return;
  ProfScope prof("join_probe", st);
  jt_probe_kernel<K><<<grid_for_n(m), JT_NT, 0, st>>>(
      q, m, t.tab.as<unsigned long long>(), t.nslots - 1, 64 - t.bits,
      t.side.as<unsigned long long>(), vals_out, found);
  FHIP_CHECK(hipGetLastError());
}

// Probes q and writes the pairs (lidx of the query, looked-up value) of its
// hits in query order (at most capacity of them); returns the hit count
// (synchronises the stream).
template <class K>
size_t jt_probe_pairs(const JoinTable& t, const K* q, const size_t* lidx, size_t m, size_t* lo,
                      size_t* ro, size_t capacity, hipStream_t st) {
  if (!m) return 0;
  const size_t tiles = ceil_div(m, size_t(JP_NT) * JP_ITEMS);
  FHIP_REQUIRE(tiles < (size_t(1) << 31), FHIP_ERR_TOO_LARGE, "equi_join: too many probe rows");
  DBuf ws((tiles + 2) * sizeof(unsigned long long));
  unsigned long long* state = ws.as<unsigned long long>();
  FHIP_CHECK(hipMemsetAsync(state, 0, (tiles + 2) * sizeof(unsigned long long), st));
  {
    ProfScope prof("join_probe", st);
    jt_probe_compact_kernel<K><<<unsigned(tiles), JP_NT, 0, st>>>(
        q, m, t.tab.as<unsigned long long>(), t.nslots - 1, 64 - t.bits,
        t.side.as<unsigned long long>(), lidx, lo, ro, capacity, state,
        reinterpret_cast<unsigned*>(state + tiles), state + tiles + 1);
    FHIP_CHECK(hipGetLastError());
  }
  unsigned long long total = 0;
  FHIP_CHECK(hipMemcpyAsync(&total, state + tiles + 1, sizeof(total), hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  return size_t(total);
}

// Builds the table over (keys, vals) and probes q with it; returns false
// (without probing) when the keys repeat.
template <class K>
bool jt_build_probe(const K* keys, const unsigned long long* vals, size_t n, const K* q, size_t m,
                    unsigned long long* vals_out, uint8_t* found, hipStream_t st) {
  JoinTable t;
  jt_build<K>(t, keys, vals, n, st);
  if (!t.unique) return false;
  jt_probe<K>(t, q, m, vals_out, found, st);
  return true;
}

// ---- star join ---------------------------------------------------------------
constexpr int SJ_MAXD = 8;

template <class K>
__global__ __launch_bounds__(256) void sj_gather_kernel(const K* __restrict__ col,
                                                        const size_t* __restrict__ idx, size_t n,
                                                        K* __restrict__ out) {
  const size_t stride = size_t(gridDim.x) * 256;
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride) out[i] = col[idx[i]];
}

__global__ __launch_bounds__(256) void sj_and_kernel(uint8_t* __restrict__ acc,
                                                     const uint8_t* __restrict__ f, size_t n) {
  const size_t stride = size_t(gridDim.x) * 256;
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride) acc[i] &= f[i];
}

// A dimension whose unique keys fill their range (at most twice the rows,
// at most 2^24 keys) and whose row ids are ids[0] + i is probed through an
// array indexed by key - min that holds each key's row (4 bytes a key: at
// most 64 MB, 4 MB for 2^20 keys, so the probes hit the cache).
constexpr unsigned SJ_EMPTY = ~0u;

template <class K>
__device__ __forceinline__ unsigned long long sj_ordered(K k) {
  if constexpr (std::is_signed<K>::value) return (unsigned long long)(long long)k ^ (1ull << 63);
  else return (unsigned long long)k;
}

// st[0] min, st[1] max of the ordered keys, st[2] != 0 when the ids are
// not ids[0] + i; the caller sets st to {~0, 0, 0} first
template <class K>
__global__ __launch_bounds__(256) void sj_stats_kernel(const K* __restrict__ keys,
                                                       const size_t* __restrict__ ids, size_t n,
                                                       unsigned long long* __restrict__ st) {
  unsigned long long mn = ~0ull, mx = 0ull;
  bool bad = false;
  const size_t id0 = ids[0];
  const size_t stride = size_t(gridDim.x) * 256;
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride) {
    const unsigned long long u = sj_ordered(keys[i]);
    mn = u < mn ? u : mn;
    mx = u > mx ? u : mx;
    if (ids[i] != id0 + i) bad = true;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long a = __shfl_xor(mn, o, 64), b = __shfl_xor(mx, o, 64);
    mn = a < mn ? a : mn;
    mx = b > mx ? b : mx;
  }
  const bool anybad = __ballot(bad) != 0ull;
  if (__lane_id() == 0) {
    atomicMin(&st[0], mn);
    atomicMax(&st[1], mx);
    if (anybad) atomicOr(&st[2], 1ull);
  }
}

__global__ void sj_init_kernel(unsigned long long* __restrict__ st, unsigned* __restrict__ tab,
                               size_t R) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < R; i += stride)
    tab[i] = SJ_EMPTY;
  if (blockIdx.x == 0 && threadIdx.x == 0) st[3] = 0ull;
}

// st[3] != 0 when a key repeats
template <class K>
__global__ __launch_bounds__(256) void sj_build_kernel(const K* __restrict__ keys, size_t n,
                                                       unsigned long long mn,
                                                       unsigned* __restrict__ tab,
                                                       unsigned long long* __restrict__ st) {
  const size_t stride = size_t(gridDim.x) * 256;
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride)
    if (atomicCAS(&tab[sj_ordered(keys[i]) - mn], SJ_EMPTY, unsigned(i)) != SJ_EMPTY)
      atomicOr(&st[3], 1ull);
}

template <class K>
__global__ __launch_bounds__(256) void sj_probe_kernel(const K* __restrict__ q, size_t m,
                                                       unsigned long long mn, unsigned long long R,
                                                       const unsigned* __restrict__ tab,
                                                       const size_t* __restrict__ ids,
                                                       unsigned long long* __restrict__ vals,
                                                       uint8_t* __restrict__ found) {
  const size_t id0 = ids[0];
  const size_t stride = size_t(gridDim.x) * 256;
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < m; i += stride) {
    const unsigned long long o = sj_ordered(q[i]) - mn;  // wraps below mn
    const unsigned pos = o < R ? tab[o] : SJ_EMPTY;
    vals[i] = pos != SJ_EMPTY ? id0 + pos : 0ull;
    found[i] = pos != SJ_EMPTY ? 1 : 0;
  }
}

// The direct-array build and probe of one dimension; returns -1 when it
// does not apply (then the caller uses the hash table), else 1 for unique
// keys and 0 for a repeated key.
template <class K>
int sj_direct(const K* keys, const size_t* ids, size_t n, const K* q, size_t m,
              unsigned long long* vals, uint8_t* found, hipStream_t st) {
  if (n < (size_t(1) << 16) || n > (size_t(1) << 24) || tune_get("join.star_direct", 1) == 0)
    return -1;
  DBuf sb(4 * sizeof(unsigned long long));
  unsigned long long* stv = sb.as<unsigned long long>();
  const unsigned long long init[3] = {~0ull, 0ull, 0ull};
  unsigned long long* h = static_cast<unsigned long long*>(pinned_scratch(0, 4 * 8));
  std::memcpy(h, init, sizeof(init));
  FHIP_CHECK(hipMemcpyAsync(stv, h, sizeof(init), hipMemcpyHostToDevice, st));
  sj_stats_kernel<K><<<grid_for_n(n), 256, 0, st>>>(keys, ids, n, stv);
  FHIP_CHECK(hipGetLastError());
  FHIP_CHECK(hipMemcpyAsync(h, stv, 3 * 8, hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  const unsigned long long mn = h[0], mx = h[1];
  if (h[2] != 0 || mx < mn || mx - mn >= 2 * (unsigned long long)n) return -1;
  const unsigned long long R = mx - mn + 1;
  DBuf tab(size_t(R) * sizeof(unsigned));
  sj_init_kernel<<<grid_for_n(size_t(R)), 256, 0, st>>>(stv, tab.as<unsigned>(), size_t(R));
  sj_build_kernel<K><<<grid_for_n(n), 256, 0, st>>>(keys, n, mn, tab.as<unsigned>(), stv);
  FHIP_CHECK(hipGetLastError());
  FHIP_CHECK(hipMemcpyAsync(h + 3, stv + 3, 8, hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  if (h[3] != 0) return 0;
  sj_probe_kernel<K><<<grid_for_n(m), 256, 0, st>>>(q, m, mn, R, tab.as<unsigned>(), ids, vals,
                                                    found);
  FHIP_CHECK(hipGetLastError());
  return 1;
}

struct StarWriter {
  const size_t* lidx;  // null: the left rows are 0..nl-1
  size_t* lo;
  const uint64_t* vals[SJ_MAXD];
  size_t* ro[SJ_MAXD];
  int nd;
  __device__ void operator()(size_t pos, size_t i) const {
    lo[pos] = lidx ? lidx[i] : i;
    for (int d = 0; d < nd; d++) ro[d][pos] = vals[d][i];
  }
};

}  // namespace

// Inner join, or LEFT outer join when only_left is given (outer_equi_join,
// join.hpp:328-379): then only_left (nl entries) receives the left row ids
// without a match and *n_only their count.
// inplace: on the sort-merge path the (device) inputs are sorted in place
// with their indices, as the reference's VE body leaves them.
size_t equi_join(int ktype, const void* left, const size_t* lidx, size_t nl, const void* right,
                 const size_t* ridx, size_t nr, int strategy, size_t* lo, size_t* ro,
                 size_t capacity, int* used_hash, size_t* only_left, size_t* n_only,
                 hipStream_t st, bool inplace = false) {
  if (n_only) *n_only = 0;
  FHIP_REQUIRE(strategy >= FHIP_JOIN_AUTO && strategy <= FHIP_JOIN_AUTO_UNIQUE, FHIP_ERR_ARG,
               "equi_join: bad strategy");
  if (used_hash) *used_hash = 0;
  bool dup_seen = false;
  // AUTO_UNIQUE: AUTO for a build side the caller expects unique (the
  // learned model's prediction); the table build alone decides
  if (strategy == FHIP_JOIN_AUTO_UNIQUE) strategy = FHIP_JOIN_AUTO;
  else if (strategy == FHIP_JOIN_AUTO) {
    dispatch_type(ktype, [&](auto z) {
      using K = decltype(z);
      // integer keys only: for floating keys the reference's uniqueness
      // test compares with == and != (NaN, -0.0), which a sample cannot
      // reproduce
      if constexpr (std::is_integral<K>::value)
        dup_seen = sample_has_duplicate<K>(ktype, (const K*)right, nr, st);
    });
  }
  if (strategy != FHIP_JOIN_SORTMERGE && !dup_seen) {
    // the probe's hits in left order, then the left rows without a partner
    auto emit = [&](DBuf& vals, DBuf& found) {
      ProfScope prof("join_compact", st);
      const size_t r = compact_apply(HitPred{found.as<uint8_t>()},
                                     HitWriter{lidx, vals.as<uint64_t>(), lo, ro}, nl, capacity,
                                     st);
      // left_idx[missed], missed ascending (join.hpp:340-350)
      if (only_left)
        *n_only = compact(MissPred{found.as<uint8_t>()}, RowIdMap{lidx}, nl, only_left, st);
      return r;
    };
    bool uniq = true, done = false;
    size_t r = 0;
    dispatch_type(ktype, [&](auto z) {
      using K = decltype(z);
      if constexpr (std::is_integral<K>::value) {
        // a unique key column that fills its range: the partitioned direct
        // table (densejoin.hip); else the open-addressing table
        const int dres = dense_join_pairs(ktype, left, lidx, nl, right, ridx, nr, lo, ro,
                                          capacity, &r, only_left, n_only, st);
        if (dres != DJ_NA) {
          uniq = dres == DJ_DONE;
          done = true;
          return;
        }
        JoinTable t;
        jt_build<K>(t, (const K*)right, reinterpret_cast<const unsigned long long*>(ridx), nr, st);
        uniq = t.unique;
        if (uniq && !only_left) {
          // inner join: the probe writes the pairs itself, in left order
          r = jt_probe_pairs<K>(t, (const K*)left, lidx, nl, lo, ro, capacity, st);
        } else if (uniq) {
          DBuf vals(nl * sizeof(uint64_t)), found(nl);
          jt_probe<K>(t, (const K*)left, nl, vals.as<unsigned long long>(), found.as<uint8_t>(),
                      st);
          r = emit(vals, found);
        }
        done = true;
      }
    });
    if (!done) {
      int u = 1;
      fhip_ht* h = ht_build(ktype, right, reinterpret_cast<const uint64_t*>(ridx), nr, true, &u, st);
      uniq = u != 0;
      if (uniq) {
        try {
          DBuf vals(nl * sizeof(uint64_t)), found(nl);
          ht_probe(h, left, nl, vals.as<uint64_t>(), found.as<uint8_t>(), st);
          r = emit(vals, found);
        } catch (...) {
          ht_destroy(h);
          throw;
        }
      }
      ht_destroy(h);
    }
    if (uniq) {
      if (used_hash) *used_hash = 1;
      return r;
    }
    FHIP_REQUIRE(strategy == FHIP_JOIN_AUTO, FHIP_ERR_ARG,
                 "equi_join: hash strategy needs unique right keys");
  }
  size_t r = 0;
  dispatch_type(ktype, [&](auto z) {
    using K = decltype(z);
    r = sort_merge_t<K>(ktype, (const K*)left, lidx, nl, (const K*)right, ridx, nr, lo, ro,
                        capacity, only_left, n_only, st, inplace);
  });
  return r;
}

}  // namespace fhip

extern "C" int fhip_dev_equi_join(int ktype, const void* left, const size_t* left_idx, size_t nl,
                                  const void* right, const size_t* right_idx, size_t nr,
                                  int strategy, size_t* left_out, size_t* right_out,
                                  size_t capacity, size_t* nout, int* used_hash, void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(nout, FHIP_ERR_ARG, "equi_join: null nout");
    FHIP_REQUIRE((nl == 0 || (left && left_idx)) && (nr == 0 || (right && right_idx)),
                 FHIP_ERR_ARG, "equi_join: null input");
    FHIP_REQUIRE(capacity == 0 || (left_out && right_out), FHIP_ERR_ARG, "equi_join: null output");
    size_t r = fhip::equi_join(ktype, left, left_idx, nl, right, right_idx, nr, strategy, left_out,
                               right_out, capacity, used_hash, nullptr, nullptr,
                               fhip::as_stream(stream));
    *nout = r;
    FHIP_REQUIRE(r <= capacity, FHIP_ERR_TOO_LARGE,
                 "equi_join: output capacity " + std::to_string(capacity) + " < pairs " +
                     std::to_string(r));
  });
}

extern "C" int fhip_dev_outer_equi_join(int ktype, const void* left, const size_t* left_idx,
                                        size_t nl, const void* right, const size_t* right_idx,
                                        size_t nr, int strategy, size_t* left_out,
                                        size_t* right_out, size_t capacity, size_t* nout,
                                        size_t* only_left, size_t* n_only_left, int* used_hash,
                                        void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(nout && n_only_left, FHIP_ERR_ARG, "outer_equi_join: null count output");
    FHIP_REQUIRE((nl == 0 || (left && left_idx && only_left)) &&
                     (nr == 0 || (right && right_idx)),
                 FHIP_ERR_ARG, "outer_equi_join: null input");
    FHIP_REQUIRE(capacity == 0 || (left_out && right_out), FHIP_ERR_ARG,
                 "outer_equi_join: null output");
    size_t r = fhip::equi_join(ktype, left, left_idx, nl, right, right_idx, nr, strategy, left_out,
                               right_out, capacity, used_hash, only_left, n_only_left,
                               fhip::as_stream(stream));
    *nout = r;
    FHIP_REQUIRE(r <= capacity, FHIP_ERR_TOO_LARGE,
                 "outer_equi_join: output capacity " + std::to_string(capacity) + " < pairs " +
                     std::to_string(r));
  });
}

namespace fhip {

// star_join on one rank (dftable.cc:295-333 -> star_join_eq,
// dfcolumn_join.cc:561-605 -> unique_equi_join2, join.hpp:103-117): for
// every dimension d a unique_hashtable over its keys with the right row ids
// as values ("star_join: right column is not unique" otherwise) is probed
// with the left keys; the left rows found in every dimension survive, in
// left order, with the looked-up right ids.  The reference probes only the
// survivors of the earlier dimensions and shrinks the outputs after each
// (shrink_missed_inplace), which keeps the same rows in the same order with
// the same looked-up ids.
size_t star_join(int nd, const int* ktypes, const void* const* left_cols, const size_t* lidx,
                 size_t nl, const void* const* right_cols, const size_t* const* ridx,
                 const size_t* nr, size_t* lo, size_t* const* ro, hipStream_t st) {
  FHIP_REQUIRE(nd >= 1 && nd <= SJ_MAXD, FHIP_ERR_ARG,
               "star_join: 1 to " + std::to_string(SJ_MAXD) + " dimension tables");
  for (int d = 0; d < nd; d++) type_bytes(ktypes[d]);
  if (nl == 0) return 0;
  std::vector<DBuf> vals(nd);
  DBuf found(nl), fd(nl), keys;
  StarWriter w{};
  w.lidx = lidx;
  w.lo = lo;
  w.nd = nd;
  for (int d = 0; d < nd; d++) {
    const void* q = left_cols[d];
    if (lidx) {
      keys.reset(nl * size_t(type_bytes(ktypes[d])));
      dispatch_type(ktypes[d], [&](auto z) {
        using K = decltype(z);
        sj_gather_kernel<K><<<grid_for_n(nl), 256, 0, st>>>((const K*)left_cols[d], lidx, nl,
                                                            keys.as<K>());
      });
      FHIP_CHECK(hipGetLastError());
      q = keys.get();
    }
    vals[d].reset(nl * sizeof(uint64_t));
    uint8_t* f = d == 0 ? found.as<uint8_t>() : fd.as<uint8_t>();
    bool done = false;
    dispatch_type(ktypes[d], [&](auto z) {
      using K = decltype(z);
      if constexpr (std::is_integral<K>::value) {
        int direct = -1;
        {
          ProfScope prof("join_probe", st);
          direct = sj_direct<K>((const K*)right_cols[d], ridx[d], nr[d], (const K*)q, nl,
                                vals[d].as<unsigned long long>(), f, st);
        }
        const bool uniq =
            direct >= 0 ? direct == 1
                        : jt_build_probe<K>((const K*)right_cols[d],
                                            reinterpret_cast<const unsigned long long*>(ridx[d]),
                                            nr[d], (const K*)q, nl,
                                            vals[d].as<unsigned long long>(), f, st);
        FHIP_REQUIRE(uniq, FHIP_ERR_ARG, "star_join: right column is not unique");
        done = true;
      }
    });
    if (!done) {
      int uniq = 1;
      fhip_ht* h = ht_build(ktypes[d], right_cols[d], reinterpret_cast<const uint64_t*>(ridx[d]),
                            nr[d], true, &uniq, st);
      try {
        FHIP_REQUIRE(uniq, FHIP_ERR_ARG, "star_join: right column is not unique");
        ht_probe(h, q, nl, vals[d].as<uint64_t>(), f, st);
      } catch (...) {
        ht_destroy(h);
        throw;
      }
      ht_destroy(h);
    }
    if (d > 0) {
      sj_and_kernel<<<grid_for_n(nl), 256, 0, st>>>(found.as<uint8_t>(), f, nl);
      FHIP_CHECK(hipGetLastError());
    }
    w.vals[d] = vals[d].as<uint64_t>();
    w.ro[d] = ro[d];
  }
  ProfScope prof("join_compact", st);
  return compact_apply(HitPred{found.as<uint8_t>()}, w, nl, nl, st);
}

}  // namespace fhip

extern "C" int fhip_dev_star_join(int ndims, const int* ktypes, const void* const* left_cols,
                                  const size_t* left_idx, size_t nl,
                                  const void* const* right_cols,
                                  const size_t* const* right_idx, const size_t* nr,
                                  size_t* left_out, size_t* const* right_out, size_t* nout,
                                  void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(nout && ktypes && left_cols && right_cols && right_idx && nr && right_out,
                 FHIP_ERR_ARG, "star_join: null pointer");
    FHIP_REQUIRE(nl == 0 || left_out, FHIP_ERR_ARG, "star_join: null output");
    for (int d = 0; d < ndims; d++) {
      FHIP_REQUIRE(nl == 0 || (left_cols[d] && right_out[d]), FHIP_ERR_ARG,
                   "star_join: null left column or output");
      FHIP_REQUIRE(nr[d] == 0 || (right_cols[d] && right_idx[d]), FHIP_ERR_ARG,
                   "star_join: null right column");
    }
    *nout = fhip::star_join(ndims, ktypes, left_cols, left_idx, nl, right_cols, right_idx, nr,
                            left_out, right_out, fhip::as_stream(stream));
  });
}
