// Reusable device building blocks: stable index compaction (one pass with a
// decoupled look-back between tiles, and a two-pass form for graph capture)
// and the single-workgroup scan of per-tile counts.  Header-only so every
// kernel file can instantiate them with its own predicates.
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

// ---- one-pass compaction ---------------------------------------------------
// Each workgroup takes the next tile by ticket (so every earlier tile has
// started and the look-back below always ends) and evaluates the predicate
// on its NT x ITEMS indices once, striped: in round j thread t takes index
// j * NT + t (coalesced), and per-wave ballots give the counts and every
// lane's rank.  The tile publishes its count; wave 0 then reads the words of
// up to 64 earlier tiles at once (aggregate or inclusive prefix, flag in the
// top two bits) until it reaches an inclusive prefix, and the selected
// entries are written in index order.  The input is read once: half the
// traffic of count-then-write.  (A blocked layout -- ITEMS consecutive
// indices per thread -- and a look-back by the whole workgroup measured
// slower.)
// Tile shapes (threads, items per thread), chosen by the tunable
// "compact.shape"; 1024 x 16 is the fastest measured at 2^24 and 2^28 rows.
constexpr int kCompactShapes[][2] = {{256, 16}, {1024, 16}};
constexpr int kCompactDefaultShape = 1;
constexpr unsigned long long CP_FLAG_A = 1ull << 62;  // tile count only
constexpr unsigned long long CP_FLAG_P = 2ull << 62;  // inclusive prefix
constexpr unsigned long long CP_VALUE = (1ull << 62) - 1;

__device__ __forceinline__ unsigned lane_rank(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi(unsigned(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(unsigned(mask), 0u));
}

// Publishes the tile's count `run` and returns the count of all earlier
// tiles; called by all 64 lanes of one wave.
__device__ __forceinline__ unsigned long long compact_lookback(unsigned long long* tile_state,
                                                               unsigned tile,
                                                               unsigned long long run) {
  const unsigned lane = __lane_id();
  unsigned long long* mine = tile_state + tile;
  if (lane == 0)
    __hip_atomic_store(mine, (tile == 0 ? CP_FLAG_P : CP_FLAG_A) | run, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  unsigned long long excl = 0;
  if (tile == 0) return 0;
  long long p = (long long)tile - 1;  // lane l looks at tile p - l
  while (true) {
    const unsigned long long v =
        (p - (long long)lane >= 0)
            ? __hip_atomic_load(tile_state + (p - lane), __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT)
            : CP_FLAG_P;  // before tile 0: an empty inclusive prefix
    const unsigned long long flag = v & ~CP_VALUE;
    const unsigned long long not_ready = __ballot(flag == 0);
    const unsigned long long prefix = __ballot(flag == CP_FLAG_P);
    // lanes up to the first inclusive prefix, if no unpublished tile comes
    // before it; else the published lanes before the first gap
    const int first_gap = not_ready ? __ffsll((long long)not_ready) - 1 : 64;
    const int first_p = prefix ? __ffsll((long long)prefix) - 1 : 64;
    const bool done = first_p < first_gap;
    const int take = done ? first_p + 1 : first_gap;
    unsigned long long part = (int(lane) < take) ? (v & CP_VALUE) : 0ull;
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
    excl += part;
    if (done) break;
    if (take == 0) __builtin_amdgcn_s_sleep(1);
    p -= take;
  }
  if (lane == 0)
    __hip_atomic_store(mine, CP_FLAG_P | (excl + run), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  return excl;
}

// Writer(pos, i) is called for the selected indices with pos < capacity;
// tile_state (nb words), ticket and total start at zero.
template <int NT, int ITEMS, class Pred, class Writer>
__global__ __launch_bounds__(NT) void compact_striped_kernel(
    Pred pred, Writer writer, size_t n, size_t capacity, unsigned long long* tile_state,
    unsigned* ticket, unsigned long long* total) {
  constexpr int WAVES = NT / 64;
  constexpr int COUNTS = ITEMS * WAVES;  // (round, wave) counts, round-major
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
  const size_t base = size_t(tile) * (NT * ITEMS);
  unsigned long long mask[ITEMS];
  if (base + size_t(NT * ITEMS) <= n) {
    // full tile: no bounds test, so all ITEMS predicate loads issue together
    bool f[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; j++) f[j] = pred(base + size_t(j) * NT + threadIdx.x);
#pragma unroll
    for (int j = 0; j < ITEMS; j++) mask[j] = __ballot(f[j]);
  } else {
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
      const size_t i = base + size_t(j) * NT + threadIdx.x;
      mask[j] = __ballot(i < n && pred(i));
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < ITEMS; j++) s_cnt[j * WAVES + wave] = unsigned(__popcll(mask[j]));
  }
  __syncthreads();
  if (wave == 0) {
    // exclusive scan of the counts: PER_LANE consecutive ones per lane
    unsigned c[PER_LANE];
    unsigned sum = 0;
#pragma unroll
    for (int q = 0; q < PER_LANE; q++) {
      c[q] = s_cnt[lane * PER_LANE + q];
      sum += c[q];
    }
    const unsigned incl = wave_incl_scan(sum);
    const unsigned run = __shfl(incl, 63, 64);
    unsigned e = incl - sum;
#pragma unroll
    for (int q = 0; q < PER_LANE; q++) {
      s_cnt[lane * PER_LANE + q] = e;
      e += c[q];
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
  for (int j = 0; j < ITEMS; j++) {
    if ((mask[j] >> lane) & 1ull) {
      const size_t pos = size_t(excl + s_cnt[j * WAVES + wave] + lane_rank(mask[j]));
      if (pos < capacity) writer(pos, base + size_t(j) * NT + threadIdx.x);
    }
  }
}

template <class Map, class OutT>
struct MapWriter {
  Map map;
  OutT* out;
  __device__ void operator()(size_t pos, size_t i) const { out[pos] = map(i); }
};

// Runs the one-pass compaction and returns the number of selected indices
// (synchronises the stream).
template <int NT, int ITEMS, class Pred, class Writer>
void launch_compact_onepass(Pred pred, Writer writer, size_t n, size_t capacity,
                            unsigned long long* state, size_t nb, hipStream_t st) {
  compact_striped_kernel<NT, ITEMS><<<unsigned(nb), NT, 0, st>>>(
      pred, writer, n, capacity, state, reinterpret_cast<unsigned*>(state + nb), state + nb + 1);
}

template <class Pred, class Writer>
size_t compact_onepass(Pred pred, Writer writer, size_t n, size_t capacity, hipStream_t st) {
  if (n == 0) return 0;
  long long shape = tune_get("compact.shape", kCompactDefaultShape);
  if (shape < 0 || shape >= (long long)(sizeof(kCompactShapes) / sizeof(kCompactShapes[0])))
    shape = kCompactDefaultShape;
  const size_t tile = size_t(kCompactShapes[shape][0]) * size_t(kCompactShapes[shape][1]);
  const size_t nb = ceil_div(n, tile);
  DBuf ws((nb + 2) * sizeof(unsigned long long));
  unsigned long long* state = ws.as<unsigned long long>();
  FHIP_CHECK(hipMemsetAsync(state, 0, (nb + 2) * sizeof(unsigned long long), st));
  if (shape == 0)
    launch_compact_onepass<256, 16>(pred, writer, n, capacity, state, nb, st);
  else
    launch_compact_onepass<1024, 16>(pred, writer, n, capacity, state, nb, st);
  FHIP_CHECK(hipGetLastError());
  unsigned long long total = 0;
  FHIP_CHECK(hipMemcpyAsync(&total, state + nb + 1, sizeof(total), hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  return size_t(total);
}

// Stable compaction; returns the number of selected indices (synchronises
// the stream).  `out` must hold every selected entry.
template <class Pred, class Map, class OutT>
size_t compact(Pred pred, Map map, size_t n, OutT* out, hipStream_t st) {
  return compact_onepass(pred, MapWriter<Map, OutT>{map, out}, n, ~size_t(0), st);
}

template <class Pred>
size_t compact_indices(Pred pred, size_t n, size_t* out, hipStream_t st) {
  return compact(pred, IdentityMap{}, n, out, st);
}

// Counts the selected indices (synchronises) and runs the writer for the
// first `capacity` of them (pos < capacity) in the same pass.  Returns the
// count; a count above `capacity` means the output was cut short.
template <class Pred, class Writer>
size_t compact_apply(Pred pred, Writer writer, size_t n, size_t capacity, hipStream_t st) {
  return compact_onepass(pred, writer, n, capacity, st);
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

// Rows of a tile-local partition: tile t (TILE rows) holds its rows grouped
// by part at t * TILE + starts[p * ntiles + t] .. t * TILE + starts[(p + 1) *
// ntiles + t] - 1.  Calls f(pos[U], in[U]) for the rows of part p in tiles
// [t0, t1), U rows per lane at a time; pos is 0 where in is false, so that
// loads through it need no branch.  A wave takes 64 tiles at once (lane l
// reads tile g + l's segment bounds) and spreads their rows over its lanes:
// row r (the r-th of the 64 segments together) belongs to the first tile
// whose inclusive sum of lengths exceeds r.  The rows of a lane only grow, so each lane keeps a
// cursor on that tile and moves it forward, the whole wave one step at a
// time (about two steps per 64 rows).
//
// SEARCH: a lane finds its tile by a binary search over the 64 inclusive
// sums instead (six shuffles a row, whatever the segment lengths), for
// segments much shorter than a wave.
//
// PosT: the type of the positions (unsigned when every position of the
// tiles fits 32 bits, which saves a register per row in flight).
template <int TILE, int U, bool SEARCH = false, class PosT = size_t, class F>
__device__ __forceinline__ void segment_rows(const unsigned short* __restrict__ starts,
                                            size_t ntiles, int p, size_t t0, size_t t1,
                                            unsigned wave, unsigned nwaves, F&& f) {
  const unsigned lane = __lane_id();
  const unsigned short* sp = starts + size_t(p) * ntiles;
  const unsigned short* ep = sp + ntiles;
  for (size_t g = t0 + size_t(wave) * 64; g < t1; g += size_t(nwaves) * 64) {
    const size_t tile = g + lane;
    unsigned s = 0, len = 0;
    if (tile < t1) {
      s = sp[tile];
      len = unsigned(ep[tile]) - s;
    }
    const unsigned incl = wave_incl_scan(len);
    const unsigned total = __shfl(incl, 63, 64);
    // row r of the lane's tile sits at tile * TILE + (s - excl) + r
    const unsigned shift = s - (incl - len);
    unsigned cur = 0;
    for (unsigned r0 = 0; r0 < total; r0 += 64 * U) {
      PosT pos[U];
      bool in[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const unsigned r = r0 + unsigned(u) * 64 + lane;
        if constexpr (SEARCH) {
          // lanes before the tile of row r: incl <= r (every lane shuffles)
          unsigned idx = 0;
#pragma unroll
          for (unsigned step = 32; step > 0; step >>= 1) {
            const unsigned nx = unsigned(__shfl(incl, int(idx + step - 1), 64));
            if (nx <= r) idx += step;
          }
          cur = idx < 63 ? idx : 63u;
        } else {
          while (true) {
            // every lane takes part in the shuffle (a lane left out of it
            // would hand 0 to the lanes reading from it)
            const unsigned nx = unsigned(__shfl(incl, int(cur), 64));
            const bool adv = cur < 63 && nx <= r;
            if (!__any(adv)) break;
            cur += adv ? 1u : 0u;
          }
        }
        in[u] = r < total;
        const unsigned sh = __shfl(shift, int(cur), 64);
        pos[u] = in[u] ? PosT((g + cur) * TILE + unsigned(sh + r)) : PosT(0);
      }
      f(pos, in);
    }
  }
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
