// Hash join on a dense unique integer key column (the build side of
// equi_join, dataframe/join.hpp:304-326: a unique_hashtable over RIGHT with
// the right row ids as values, probed by LEFT in order).  When RIGHT's keys
// fill a range [mn, mn + R) with R at most join.dense_ratio x the rows (a
// primary-key or surrogate-id column), the table is an array indexed by
// key - mn and the join is radix-partitioned on the high bits of that
// offset, so that each part's table slice (1 MB by default) stays in one
// XCD's L2 while it is probed:
//
//   stats   one pass over RIGHT: key range, and whether the right row ids
//           are ridx[0] + i (then the table holds 4-byte positions, else
//           the 8-byte ids);
//   part    tile-local partition of a key column (16384-row tiles): each
//           tile writes its in-range offsets grouped by part, the start of
//           every part in the tile (part-major, so the readers of one part
//           read them contiguously) and, for LEFT, where each row's offset
//           went (16 bits; out-of-range rows are marked and not written);
//   build   one workgroup per part marks every offset in an LDS bitmap (a
//           second mark is a repeated key: RIGHT is not unique, the
//           reference's is_unique_ok fails) and stores the value;
//           offsets no row took get the empty marker;
//   probe   workgroups take (part, tile chunk) pairs, blockIdx % 8 being the
//           XCD under the round-robin dispatch, so the workgroups of one
//           part share an L2; the results go next to the offsets;
//   emit    tiles in order (ticket + decoupled look-back), each row reads its
//           result back through its 16-bit position and the hits are
//           written as (left_idx, right_idx) pairs in left order.
//
// The pairs are the reference's: for every left row in order, its partner
// when RIGHT holds its key.
#include "densejoin.hpp"
#include "fhip_device.hpp"

#include <algorithm>
#include <type_traits>

namespace fhip {
namespace {

constexpr int DJ_NT = 1024;                   // partition and emit workgroups
constexpr int DJ_ITEMS = 16;                  // rows per thread of a tile
constexpr int DJ_TILE = DJ_NT * DJ_ITEMS;     // 16384: tile positions fit 16 bits
constexpr int DJ_MAXP = 2048;                 // parts at most (LDS counters)
constexpr int DJ_MAXSBITS = 18;               // offsets per part at most 2^18
constexpr unsigned short DJ_MISS = 0xFFFF;    // a row outside the key range
constexpr int DJ_PNT = 256;                   // probe workgroup
constexpr int DJ_PU = 8;                      // rows per lane in flight (probe)
constexpr int DJ_BU = 4;                      // rows per lane in flight (build)
constexpr unsigned long long DJ_MAXR = 1ull << 29;  // key range at most

struct DjStats {
  unsigned long long mn, mx;  // order-preserving images of the smallest / largest key
  unsigned long long r0;      // ridx[0]
  unsigned not_affine;        // some ridx[i] != ridx[0] + i
  unsigned has_ones;          // some ridx[i] == ~0
};

template <class K>
__device__ __forceinline__ unsigned long long dj_bits(K k) {
  if constexpr (std::is_signed<K>::value)
    return static_cast<unsigned long long>(static_cast<long long>(k));
  else
    return static_cast<unsigned long long>(k);
}

// bits -> an unsigned value with the key order (the sign bit flipped for
// signed keys)
template <class K>
__host__ __device__ __forceinline__ unsigned long long dj_ord(unsigned long long b) {
  return std::is_signed<K>::value ? (b ^ (1ull << 63)) : b;
}

__device__ __forceinline__ void dj_reduce(DjStats& s) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    s.mn = std::min(s.mn, (unsigned long long)__shfl_xor(s.mn, o, 64));
    s.mx = std::max(s.mx, (unsigned long long)__shfl_xor(s.mx, o, 64));
    s.not_affine |= __shfl_xor(s.not_affine, o, 64);
    s.has_ones |= __shfl_xor(s.has_ones, o, 64);
  }
}

// One partial per workgroup (grid-stride over the rows).
template <class K>
__global__ __launch_bounds__(256) void dj_stats_kernel(const K* __restrict__ keys,
                                                       const size_t* __restrict__ ridx, size_t n,
                                                       DjStats* __restrict__ part) {
  const unsigned long long r0 = ridx[0];
  DjStats s{~0ull, 0ull, r0, 0u, 0u};
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
    const unsigned long long o = dj_ord<K>(dj_bits(keys[i]));
    const unsigned long long r = ridx[i];
    s.mn = std::min(s.mn, o);
    s.mx = std::max(s.mx, o);
    s.not_affine |= unsigned(r != r0 + i);
    s.has_ones |= unsigned(r == ~0ull);
  }
  dj_reduce(s);
  __shared__ DjStats ws[4];
  if (__lane_id() == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; w++) {
      s.mn = std::min(s.mn, ws[w].mn);
      s.mx = std::max(s.mx, ws[w].mx);
      s.not_affine |= ws[w].not_affine;
      s.has_ones |= ws[w].has_ones;
    }
    part[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(64) void dj_stats_final_kernel(const DjStats* __restrict__ part,
                                                            int g, DjStats* __restrict__ out) {
  DjStats s{~0ull, 0ull, part[0].r0, 0u, 0u};
  for (int i = threadIdx.x; i < g; i += 64) {
    const DjStats p = part[i];
    s.mn = std::min(s.mn, p.mn);
    s.mx = std::max(s.mx, p.mx);
    s.not_affine |= p.not_affine;
    s.has_ones |= p.has_ones;
  }
  dj_reduce(s);
  if (threadIdx.x == 0) *out = s;
}

// Tile-local partition of keys[0..n) by the offset o = key - mn (rows with
// o >= R are out of range) into parts of 2^sbits offsets.  Tile t writes its
// in-range offsets, grouped by part, to o_out[t * DJ_TILE ...], the start of
// part p in the tile to starts[p * ntiles + t] (p = P: the tile's in-range
// count), and
//   BUILD: the row's value next to its offset (v_out; E = 4 bytes: the row
//          position, else ridx[row]);
//   probe: dest[row] = the position the row's offset went to, DJ_MISS when
//          out of range.
// Within a part the rows come in LDS-atomic order; the readers never depend
// on it.
//   PAIRS (BUILD, 4-byte values): offset and value are written together as
//          one 8-byte pair per row to o_out (two words a row), so a reader
//          of short segments fetches one run, not two.
template <class K, class E, bool BUILD, int MAXP = DJ_MAXP, bool PAIRS = false>
__global__ __launch_bounds__(DJ_NT) void dj_part_kernel(
    const K* __restrict__ keys, size_t n, unsigned long long mn, unsigned long long R, int sbits,
    int P, size_t ntiles, const size_t* __restrict__ ridx, unsigned* __restrict__ o_out,
    E* __restrict__ v_out, unsigned short* __restrict__ dest, unsigned short* __restrict__ starts) {
  static_assert(!PAIRS || (BUILD && sizeof(E) == 4), "pairs of 4-byte offset and value");
  constexpr size_t STAGE =
      size_t(DJ_TILE) * (PAIRS ? 8 : BUILD ? sizeof(E) : sizeof(unsigned));
  __shared__ __align__(16) unsigned char s_stage[STAGE];
  __shared__ unsigned s_cnt[MAXP];
  __shared__ unsigned s_tmp[DJ_NT / 64];
  __shared__ unsigned s_total;
  unsigned* stage_o = reinterpret_cast<unsigned*>(s_stage);
  // tiles t .. t + per - 1 on one XCD (workgroups go round-robin to the
  // eight XCDs), so the start words of neighbouring tiles, which share
  // cache lines, are combined in one L2
  const size_t per = (ntiles + 7) / 8;
  const size_t t = size_t(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (t >= ntiles) return;
  const size_t base = t * DJ_TILE;
  // the tile's rows through uniform base pointers and 32-bit offsets
  const K* kt = keys + base;
  const unsigned rows = unsigned(std::min<size_t>(DJ_TILE, n - base));
  for (int p = threadIdx.x; p < MAXP; p += DJ_NT) s_cnt[p] = 0;
  // o = key - mn, or ~0u for rows out of range (R <= 2^29)
  unsigned o[DJ_ITEMS];
  if (rows == DJ_TILE) {  // a full tile: every load issued at once
    unsigned long long kb[DJ_ITEMS];
#pragma unroll
    for (int j = 0; j < DJ_ITEMS; j++) kb[j] = dj_bits(kt[unsigned(j) * DJ_NT + threadIdx.x]);
#pragma unroll
    for (int j = 0; j < DJ_ITEMS; j++) {
      const unsigned long long off = kb[j] - mn;
      o[j] = off < R ? unsigned(off) : ~0u;
    }
  } else {
#pragma unroll
    for (int j = 0; j < DJ_ITEMS; j++) {
      const unsigned i = unsigned(j) * DJ_NT + threadIdx.x;
      const unsigned long long off = i < rows ? dj_bits(kt[i]) - mn : ~0ull;
      o[j] = off < R ? unsigned(off) : ~0u;
    }
  }
  __syncthreads();
  // dr = part << 16 | rank within the part, ~0u for rows out of range
  unsigned dr[DJ_ITEMS];
#pragma unroll
  for (int j = 0; j < DJ_ITEMS; j++) {
    dr[j] = ~0u;
    if (o[j] != ~0u) {
      const unsigned d = o[j] >> sbits;
      dr[j] = (d << 16) | atomicAdd(&s_cnt[d], 1u);
    }
  }
  __syncthreads();
  {
    // exclusive scan of the MAXP counters, PER consecutive ones per thread
    constexpr int PER = MAXP / DJ_NT;
    static_assert(PER * DJ_NT == MAXP, "whole counters per thread");
    unsigned c[PER], sum = 0;
#pragma unroll
    for (int q = 0; q < PER; q++) {
      c[q] = s_cnt[PER * threadIdx.x + q];
      sum += c[q];
    }
    unsigned all;
    unsigned e = block_excl_scan<DJ_NT>(sum, s_tmp, &all);
    if (threadIdx.x == 0) s_total = all;
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int p = PER * int(threadIdx.x) + q;
      s_cnt[p] = e;
      if (p < P) starts[size_t(p) * ntiles + t] = (unsigned short)e;
      e += c[q];
    }
    if (threadIdx.x == 0) starts[size_t(P) * ntiles + t] = (unsigned short)all;
  }
  __syncthreads();
  if constexpr (PAIRS) {
    uint2* stage_p = reinterpret_cast<uint2*>(s_stage);
#pragma unroll
    for (int j = 0; j < DJ_ITEMS; j++) {
      if (dr[j] != ~0u) {
        const unsigned i = unsigned(j) * DJ_NT + threadIdx.x;
        stage_p[s_cnt[dr[j] >> 16] + (dr[j] & 0xFFFFu)] = make_uint2(o[j], unsigned(base + i));
      }
    }
    __syncthreads();
    const unsigned cnt = s_total;
    uint2* pt = reinterpret_cast<uint2*>(o_out) + base;
    for (unsigned i = threadIdx.x; i < cnt; i += DJ_NT) pt[i] = stage_p[i];
    return;
  }
  unsigned short* dt = BUILD ? nullptr : dest + base;
#pragma unroll
  for (int j = 0; j < DJ_ITEMS; j++) {
    const unsigned i = unsigned(j) * DJ_NT + threadIdx.x;
    if (dr[j] != ~0u) {
      const unsigned pos = s_cnt[dr[j] >> 16] + (dr[j] & 0xFFFFu);
      stage_o[pos] = o[j];
      if constexpr (!BUILD) dt[i] = (unsigned short)pos;
    } else if constexpr (!BUILD) {
      if (i < rows) dt[i] = DJ_MISS;
    }
  }
  __syncthreads();
  const unsigned cnt = s_total;
  unsigned* ot = o_out + base;
  for (unsigned i = threadIdx.x; i < cnt; i += DJ_NT) ot[i] = stage_o[i];
  if constexpr (BUILD) {
    E* stage_v = reinterpret_cast<E*>(s_stage);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < DJ_ITEMS; j++) {
      if (dr[j] != ~0u) {
        const unsigned i = unsigned(j) * DJ_NT + threadIdx.x;
        const unsigned pos = s_cnt[dr[j] >> 16] + (dr[j] & 0xFFFFu);
        if constexpr (sizeof(E) == 4)
          stage_v[pos] = E(base + i);
        else
          stage_v[pos] = E(ridx[base + i]);
      }
    }
    __syncthreads();
    E* vt = v_out + base;
    for (unsigned i = threadIdx.x; i < cnt; i += DJ_NT) vt[i] = stage_v[i];
  }
}

// Positions are 32-bit: the dense join takes at most 2^32 - DJ_TILE rows
// a side.
template <int U, bool SEARCH = false, class F>
__device__ __forceinline__ void dj_for_rows(const unsigned short* __restrict__ starts,
                                            size_t ntiles, int p, size_t t0, size_t t1,
                                            unsigned wave, unsigned nwaves, F&& f) {
  segment_rows<DJ_TILE, U, SEARCH, unsigned>(starts, ntiles, p, t0, t1, wave, nwaves, f);
}

// One workgroup per part: the part's offsets are marked in an LDS bitmap
// (a second mark is a repeated key, *dup is set), every row's value goes to
// table[offset], and with fill_empty the offsets no row took get `empty`.
template <class E>
__global__ __launch_bounds__(DJ_NT) void dj_build_kernel(
    const unsigned* __restrict__ o_in, const E* __restrict__ v_in,
    const unsigned short* __restrict__ starts, size_t ntiles, int sbits, unsigned long long R,
    E* __restrict__ table, E empty, int fill_empty, unsigned* __restrict__ dup) {
  __shared__ unsigned bm[(1u << DJ_MAXSBITS) / 32];
  __shared__ unsigned s_dup;
  const int p = blockIdx.x;
  const unsigned long long lo = (unsigned long long)p << sbits;
  const unsigned long long hi = std::min(R, lo + (1ull << sbits));
  const unsigned nwords = unsigned((hi - lo + 31) / 32);
  for (unsigned i = threadIdx.x; i < nwords; i += DJ_NT) bm[i] = 0;
  if (threadIdx.x == 0) s_dup = 0;
  __syncthreads();
  unsigned d = 0;
  dj_for_rows<DJ_BU>(starts, ntiles, p, 0, ntiles, threadIdx.x >> 6, DJ_NT / 64,
                     [&](const unsigned* pos, const bool* in) {
                       unsigned o[DJ_BU];
                       E v[DJ_BU];
#pragma unroll
                       for (int u = 0; u < DJ_BU; u++) {
                         o[u] = o_in[pos[u]];
                         v[u] = v_in[pos[u]];
                       }
#pragma unroll
                       for (int u = 0; u < DJ_BU; u++) {
                         if (!in[u]) continue;
                         const unsigned l = o[u] - unsigned(lo);
                         const unsigned bit = 1u << (l & 31);
                         d |= atomicOr(&bm[l >> 5], bit) & bit;
                         table[o[u]] = v[u];
                       }
                     });
  if (d) atomicOr(&s_dup, 1u);
  __syncthreads();
  if (threadIdx.x == 0 && s_dup) atomicOr(dup, 1u);
  if (fill_empty) {
    for (unsigned long long s = lo + threadIdx.x; s < hi; s += DJ_NT) {
      const unsigned l = unsigned(s - lo);
      if (!((bm[l >> 5] >> (l & 31)) & 1u)) table[s] = empty;
    }
  }
}

// One workgroup per part of at most 2^DJ_LSBITS offsets, built in LDS:
// every slot starts empty, each row claims its slot with an LDS
// compare-and-swap (a slot already taken is a repeated key: *dup is set),
// and the slice is then written out in order, empty slots included.  The
// values are 4-byte row positions, never the empty marker.
constexpr int DJ_LSBITS = 15;
constexpr int DJ_LMAXP = 4096;

__global__ __launch_bounds__(DJ_NT) void dj_lds_build_kernel(
    const uint2* __restrict__ ov, const unsigned short* __restrict__ starts, size_t ntiles,
    unsigned long long R, unsigned* __restrict__ table, unsigned* __restrict__ dup) {
  constexpr unsigned S = 1u << DJ_LSBITS;
  __shared__ __align__(16) unsigned s_slice[S];
  __shared__ unsigned s_dup;
  const int p = blockIdx.x;
  const unsigned long long lo = (unsigned long long)p << DJ_LSBITS;
  const unsigned len = unsigned(std::min<unsigned long long>(R - lo, S));
  for (unsigned i = threadIdx.x; i < S; i += DJ_NT) s_slice[i] = ~0u;
  if (threadIdx.x == 0) s_dup = 0;
  __syncthreads();
  unsigned d = 0;
  dj_for_rows<4, true>(starts, ntiles, p, 0, ntiles, threadIdx.x >> 6, DJ_NT / 64,
                       [&](const unsigned* pos, const bool* in) {
                         unsigned o[4], v[4];
#pragma unroll
                         for (int u = 0; u < 4; u++) {
                           const uint2 x = ov[pos[u]];
                           o[u] = x.x;
                           v[u] = x.y;
                         }
#pragma unroll
                         for (int u = 0; u < 4; u++) {
                           if (!in[u]) continue;
                           const unsigned l = o[u] - unsigned(lo);
                           d |= unsigned(atomicCAS(&s_slice[l], ~0u, v[u]) != ~0u);
                         }
                       });
  if (d) atomicOr(&s_dup, 1u);
  __syncthreads();
  if (threadIdx.x == 0 && s_dup) atomicOr(dup, 1u);
  unsigned* out = table + lo;
  if (len == S) {
    const uint4* src = reinterpret_cast<const uint4*>(s_slice);
    uint4* dst = reinterpret_cast<uint4*>(out);
    for (unsigned i = threadIdx.x; i < S / 4; i += DJ_NT) dst[i] = src[i];
  } else {
    for (unsigned i = threadIdx.x; i < len; i += DJ_NT) out[i] = s_slice[i];
  }
}

// What a slice pass does with each row of its part.
enum DjMode {
  DJ_PROBE = 0,    // res[pos] = table[o]
  DJ_SCATTER = 1,  // table[o] = v[pos]
  DJ_VERIFY = 2,   // *flag = 1 when table[o] != v[pos] (another row took the slot)
};

// Workgroup w takes part 8 * (g / C) + w % 8 (g = w / 8) and the tile chunk
// g % C: the C chunks of one part run on one XCD under the round-robin
// dispatch of workgroups to XCDs, so that part's table slice is read and
// written in that XCD's L2.  (Only the speed depends on the dispatch.)
template <class E, int MODE, int PU = DJ_PU>
__global__ __launch_bounds__(DJ_PNT) void dj_slice_kernel(const unsigned* __restrict__ o_in,
                                                          const unsigned short* __restrict__ starts,
                                                          size_t ntiles, int P, int C,
                                                          E* __restrict__ table,
                                                          E* __restrict__ vals,
                                                          unsigned* __restrict__ flag) {
  const unsigned w = blockIdx.x;
  const unsigned g = w >> 3;
  const int p = int((g / unsigned(C)) * 8 + (w & 7));
  if (p >= P) return;
  const size_t chunk = (ntiles + size_t(C) - 1) / size_t(C);
  const size_t t0 = size_t(g % unsigned(C)) * chunk;
  const size_t t1 = std::min(ntiles, t0 + chunk);
  if (t0 >= t1) return;
  unsigned bad = 0;
  dj_for_rows<PU>(starts, ntiles, p, t0, t1, threadIdx.x >> 6, DJ_PNT / 64,
                     [&](const unsigned* pos, const bool* in) {
                       unsigned o[PU];
#pragma unroll
                       for (int u = 0; u < PU; u++)
                         o[u] = __builtin_nontemporal_load(o_in + pos[u]);
                       if constexpr (MODE == DJ_PROBE) {
                         E v[PU];
#pragma unroll
                         for (int u = 0; u < PU; u++) v[u] = table[in[u] ? o[u] : 0u];
#pragma unroll
                         for (int u = 0; u < PU; u++)
                           if (in[u]) __builtin_nontemporal_store(v[u], vals + pos[u]);
                       } else {
                         E v[PU];
#pragma unroll
                         for (int u = 0; u < PU; u++) v[u] = __builtin_nontemporal_load(vals + pos[u]);
                         if constexpr (MODE == DJ_SCATTER) {
#pragma unroll
                           for (int u = 0; u < PU; u++)
                             if (in[u]) table[o[u]] = v[u];
                         } else {
                           E c[PU];
#pragma unroll
                           for (int u = 0; u < PU; u++) c[u] = table[in[u] ? o[u] : 0u];
#pragma unroll
                           for (int u = 0; u < PU; u++) bad |= unsigned(in[u] && c[u] != v[u]);
                         }
                       }
                     });
  if constexpr (MODE == DJ_VERIFY) {
    if (__any(bad) && __lane_id() == 0) atomicOr(flag, 1u);
  }
}

// The right row id behind a stored value.
template <class E>
__device__ __forceinline__ unsigned long long dj_rid(E v, unsigned long long rbase) {
  return sizeof(E) == 4 ? rbase + (unsigned long long)v : (unsigned long long)v;
}

// Partition tiles in order (ticket): the tile's results are first copied
// into LDS with coalesced loads, so that each row's result is read back
// through its 16-bit position from LDS, not from L2; the hits are written
// as (left_idx, right_idx) pairs after those of every earlier tile
// (decoupled look-back); pairs at positions >= capacity are counted, not
// written.  The left row ids are loaded with the positions, before the
// look-back.  Rows go in two rounds of DJ_EITEMS per thread.
// only_left (the outer join; else null): the left row ids without a
// partner, in left order.  Every tile before this one is full, so the misses
// before it are its first row minus the hits before it, and a row's misses
// within the tile are its index minus the hits before it: no second
// look-back.
constexpr int DJ_EITEMS = 8;
constexpr int DJ_EROUNDS = DJ_ITEMS / DJ_EITEMS;

template <class E>
__global__ __launch_bounds__(DJ_NT) void dj_emit_kernel(
    const unsigned short* __restrict__ dest, size_t n, const E* __restrict__ res,
    const unsigned short* __restrict__ tile_cnt, E empty, int check_empty,
    unsigned long long rbase, const size_t* __restrict__ lidx, size_t* __restrict__ lo,
    size_t* __restrict__ ro, size_t capacity, unsigned long long* tile_state, unsigned* ticket,
    unsigned long long* total, size_t* __restrict__ only_left) {
  constexpr int WAVES = DJ_NT / 64;
  constexpr int COUNTS = DJ_ITEMS * WAVES;
  constexpr int PER_LANE = COUNTS / 64;
  static_assert(COUNTS % 64 == 0, "the counts split evenly over one wave");
  __shared__ __align__(16) E s_res[DJ_TILE];
  __shared__ unsigned s_cnt[COUNTS];
  __shared__ unsigned s_tile;
  __shared__ unsigned long long s_excl;
  const unsigned lane = __lane_id();
  const unsigned wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const unsigned tile = s_tile;
  const size_t base = size_t(tile) * DJ_TILE;
  const unsigned rows = unsigned(std::min<size_t>(DJ_TILE, n - base));
  const unsigned short* dt = dest + base;
  const size_t* lt = lidx + base;
  unsigned long long hit[DJ_ITEMS];
  unsigned dpk[DJ_ITEMS / 2];  // two 16-bit positions per word
  // every position load issued at once (a clamped index, no branch)
  unsigned short dl[DJ_ITEMS];
#pragma unroll
  for (int j = 0; j < DJ_ITEMS; j++) {
    const unsigned i = unsigned(j) * DJ_NT + threadIdx.x;
    dl[j] = dt[i < rows ? i : 0u];
  }
  {
    // the tile's results into LDS, 16 bytes per load, all loads in flight
    constexpr int NV = DJ_TILE * int(sizeof(E)) / 16 / DJ_NT;
    const unsigned nvec = (unsigned(tile_cnt[tile]) * unsigned(sizeof(E)) + 15u) / 16u;
    const uint4* rt = reinterpret_cast<const uint4*>(res + base);
    uint4* sr = reinterpret_cast<uint4*>(s_res);
    uint4 rv[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) {
      const unsigned i = unsigned(k) * DJ_NT + threadIdx.x;
      rv[k] = rt[i < nvec ? i : 0u];
    }
    // (entries past the tile's results hold copies of its first ones; no
    // row reads them)
#pragma unroll
    for (int k = 0; k < NV; k++) sr[unsigned(k) * DJ_NT + threadIdx.x] = rv[k];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < DJ_ITEMS; j++) {
    const unsigned i = unsigned(j) * DJ_NT + threadIdx.x;
    const unsigned dp = i < rows ? unsigned(dl[j]) : unsigned(DJ_MISS);
    if (j & 1)
      dpk[j >> 1] |= dp << 16;
    else
      dpk[j >> 1] = dp;
    const E v = s_res[dp != DJ_MISS ? dp : 0u];
    hit[j] = __ballot(dp != DJ_MISS && !(check_empty && v == empty));
  }
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < DJ_ITEMS; j++) s_cnt[j * WAVES + wave] = unsigned(__popcll(hit[j]));
  }
  __syncthreads();
  if (wave == 0) {
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
  for (int h = 0; h < DJ_EROUNDS; h++) {
    // the left row ids of this round's hits, all loads issued together
    unsigned long long lid[DJ_EITEMS];
#pragma unroll
    for (int q = 0; q < DJ_EITEMS; q++) {
      const int j = h * DJ_EITEMS + q;
      const unsigned i = unsigned(j) * DJ_NT + threadIdx.x;
      lid[q] = lt[i < rows ? i : 0u];
    }
#pragma unroll
    for (int q = 0; q < DJ_EITEMS; q++) {
      const int j = h * DJ_EITEMS + q;
      const unsigned before = s_cnt[j * WAVES + wave] + lane_rank(hit[j]);  // hits in the tile
      if ((hit[j] >> lane) & 1ull) {
        const size_t pos = size_t(excl + before);
        if (pos < capacity) {
          const unsigned dp = (dpk[j >> 1] >> ((j & 1) * 16)) & 0xFFFFu;
          lo[pos] = lid[q];
          ro[pos] = dj_rid(s_res[dp], rbase);
        }
      } else if (only_left) {
        const unsigned i = unsigned(j) * DJ_NT + threadIdx.x;
        if (i < rows) only_left[size_t(base - excl) + (i - before)] = lid[q];
      }
    }
  }
}

struct DenseTable {
  unsigned long long mn = 0, R = 0, rbase = 0;
  int sbits = 0, P = 0, check_empty = 0;
  DBuf table;
};

// The key range and ridx shape of RIGHT; false when the dense table does not
// apply (too few rows, range too wide, ids that collide with the marker).
template <class K>
bool dj_plan(const K* right, const size_t* ridx, size_t nr, DenseTable& t, bool& wide,
             hipStream_t st) {
  const long long min_rows = tune_get("join.dense_min", 1 << 20);
  const long long ratio = std::min(64LL, tune_get("join.dense_ratio", 2));
  if (ratio <= 0 || min_rows < 0 || nr < size_t(std::max(1LL, min_rows)) ||
      nr >= size_t(0xFFFFFFFEu))
    return false;
  const int g = std::min(grid_for_n(nr, 4), 1024);
  DBuf part(size_t(g + 1) * sizeof(DjStats));
  DjStats* pp = part.as<DjStats>();
  {
    ProfScope prof("join_dense_stats", st);
    dj_stats_kernel<K><<<g, 256, 0, st>>>(right, ridx, nr, pp);
    FHIP_CHECK(hipGetLastError());
    dj_stats_final_kernel<<<1, 64, 0, st>>>(pp, g, pp + g);
    FHIP_CHECK(hipGetLastError());
  }
  DjStats* h = static_cast<DjStats*>(pinned_scratch(0, sizeof(DjStats)));
  FHIP_CHECK(hipMemcpyAsync(h, pp + g, sizeof(DjStats), hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  const DjStats s = *h;
  const unsigned long long mnb = dj_ord<K>(s.mn), mxb = dj_ord<K>(s.mx);
  const unsigned long long span = mxb - mnb;  // modular: the key bits are sign-extended
  if (span >= DJ_MAXR || span + 1 > (unsigned long long)ratio * nr) return false;
  t.mn = mnb;
  t.R = span + 1;
  wide = s.not_affine != 0;
  t.check_empty = t.R > nr;
  if (wide && t.check_empty && s.has_ones) return false;
  t.rbase = s.r0;
  int sbits = int(std::max(10LL, std::min<long long>(DJ_MAXSBITS,
                                                      tune_get("join.dense_slice_bits", 18))));
  while (sbits < DJ_MAXSBITS && ceil_div(size_t(t.R), size_t(1) << sbits) > size_t(DJ_MAXP))
    sbits++;
  t.sbits = sbits;
  t.P = int(ceil_div(size_t(t.R), size_t(1) << sbits));
  return t.P <= DJ_MAXP;
}

// Workgroups per part of a slice pass over nt tiles: join.dense_chunks, but
// at least 64 tiles for each wave of a workgroup.
inline int dj_chunks(size_t nt) {
  const long long c = tune_get("join.dense_chunks", 128);
  const long long most = std::max<long long>(1, (long long)(nt / (64 * (DJ_PNT / 64))));
  return int(std::max<long long>(1, std::min<long long>(c, most)));
}

// Builds the table over RIGHT; false when a key repeats.
template <class K, class E>
bool dj_build(const K* right, const size_t* ridx, size_t nr, DenseTable& t, hipStream_t st) {
  const size_t nt = ceil_div(nr, size_t(DJ_TILE));
  // 4-byte positions over a range of at most 2^27 offsets: parts of 2^15
  // offsets built in LDS; else the probe's parts
  const bool lds = sizeof(E) == 4 && t.R <= (unsigned long long)DJ_LMAXP << DJ_LSBITS &&
                   tune_get("join.dense_lds_build", 1) != 0;
  const int sb = lds ? DJ_LSBITS : t.sbits;
  const int P = lds ? int(ceil_div(size_t(t.R), size_t(1) << DJ_LSBITS)) : t.P;
  // (the LDS build reads offset-value pairs from o: two words a row)
  DBuf o(nt * DJ_TILE * (lds ? 8 : sizeof(unsigned))), v(lds ? 0 : nt * DJ_TILE * sizeof(E));
  DBuf starts((size_t(P) + 1) * nt * sizeof(unsigned short));
  {
    ProfScope prof("join_dense_build_part", st);
    if constexpr (sizeof(E) == 4) {
      if (lds)
        dj_part_kernel<K, E, true, DJ_LMAXP, true><<<unsigned(ceil_div(nt, 8) * 8), DJ_NT, 0, st>>>(
            right, nr, t.mn, t.R, sb, P, nt, ridx, o.as<unsigned>(), nullptr, nullptr,
            starts.as<unsigned short>());
    }
    if (!lds)
      dj_part_kernel<K, E, true><<<unsigned(ceil_div(nt, 8) * 8), DJ_NT, 0, st>>>(
          right, nr, t.mn, t.R, sb, P, nt, ridx, o.as<unsigned>(), v.as<E>(), nullptr,
          starts.as<unsigned short>());
    FHIP_CHECK(hipGetLastError());
  }
  ProfScope prof("join_dense_build", st);
  t.table.reset(size_t(t.R) * sizeof(E));
  DBuf dup(sizeof(unsigned));
  FHIP_CHECK(hipMemsetAsync(dup.get(), 0, sizeof(unsigned), st));
  if constexpr (sizeof(E) == 4) {
    if (lds) {
      dj_lds_build_kernel<<<unsigned(P), DJ_NT, 0, st>>>(
          o.as<uint2>(), starts.as<unsigned short>(), nt, t.R, t.table.as<unsigned>(),
          dup.as<unsigned>());
      FHIP_CHECK(hipGetLastError());
    } else {
      // 4-byte positions (unique per row): scatter, then check every row
      // still owns its slot (a repeated key lost one), both XCD-aware like
      // the probe
      if (t.check_empty)
        FHIP_CHECK(hipMemsetAsync(t.table.get(), 0xFF, size_t(t.R) * sizeof(E), st));
      const int C = dj_chunks(nt);
      const size_t p8 = ceil_div(size_t(P), 8) * 8;
      {
        ProfScope prof2("join_dense_scatter", st);
        dj_slice_kernel<E, DJ_SCATTER><<<unsigned(p8 * size_t(C)), DJ_PNT, 0, st>>>(
            o.as<unsigned>(), starts.as<unsigned short>(), nt, P, C, t.table.as<E>(), v.as<E>(),
            nullptr);
        FHIP_CHECK(hipGetLastError());
      }
      {
        ProfScope prof2("join_dense_verify", st);
        dj_slice_kernel<E, DJ_VERIFY><<<unsigned(p8 * size_t(C)), DJ_PNT, 0, st>>>(
            o.as<unsigned>(), starts.as<unsigned short>(), nt, P, C, t.table.as<E>(), v.as<E>(),
            dup.as<unsigned>());
        FHIP_CHECK(hipGetLastError());
      }
    }
  } else {
    // 8-byte row ids (which may repeat): a repeated key is found by one
    // workgroup per part marking its offsets in an LDS bitmap
    dj_build_kernel<E><<<unsigned(P), DJ_NT, 0, st>>>(
        o.as<unsigned>(), v.as<E>(), starts.as<unsigned short>(), nt, sb, t.R, t.table.as<E>(),
        E(~E(0)), t.check_empty, dup.as<unsigned>());
    FHIP_CHECK(hipGetLastError());
  }
  unsigned* h = static_cast<unsigned*>(pinned_scratch(0, sizeof(unsigned)));
  FHIP_CHECK(hipMemcpyAsync(h, dup.get(), sizeof(unsigned), hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  return *h == 0;
}

template <class K, class E>
void dj_probe(const DenseTable& t, const K* left, size_t nl, DBuf& dest, DBuf& res,
              DBuf& starts, hipStream_t st) {
  const size_t nt = ceil_div(nl, size_t(DJ_TILE));
  DBuf o(nt * DJ_TILE * sizeof(unsigned));
  starts.reset((size_t(t.P) + 1) * nt * sizeof(unsigned short));
  dest.reset(nl * sizeof(unsigned short));
  res.reset(nt * DJ_TILE * sizeof(E));
  {
    ProfScope prof("join_dense_part", st);
    dj_part_kernel<K, E, false><<<unsigned(ceil_div(nt, 8) * 8), DJ_NT, 0, st>>>(
        left, nl, t.mn, t.R, t.sbits, t.P, nt, nullptr, o.as<unsigned>(), nullptr,
        dest.as<unsigned short>(), starts.as<unsigned short>());
    FHIP_CHECK(hipGetLastError());
  }
  const int C = dj_chunks(nt);
  const size_t p8 = ceil_div(size_t(t.P), 8) * 8;
  {
    ProfScope prof("join_dense_probe", st);
    // rows per lane in flight: join.dense_probe_u (4, 8, 16 or 32)
    // (measured for 2^29 probe rows: 4.32, 3.67-3.80, 3.36-3.65 and 3.25 ms
    // for 4, 8, 16 and 32: the probe waits on its L2 reads)
    const long long pu = tune_get("join.dense_probe_u", 32);
    const unsigned g = unsigned(p8 * size_t(C));
    if (pu <= 4)
      dj_slice_kernel<E, DJ_PROBE, 4><<<g, DJ_PNT, 0, st>>>(
          o.as<unsigned>(), starts.as<unsigned short>(), nt, t.P, C, t.table.as<E>(),
          res.as<E>(), nullptr);
    else if (pu >= 32)
      dj_slice_kernel<E, DJ_PROBE, 32><<<g, DJ_PNT, 0, st>>>(
          o.as<unsigned>(), starts.as<unsigned short>(), nt, t.P, C, t.table.as<E>(),
          res.as<E>(), nullptr);
    else if (pu >= 16)
      dj_slice_kernel<E, DJ_PROBE, 16><<<g, DJ_PNT, 0, st>>>(
          o.as<unsigned>(), starts.as<unsigned short>(), nt, t.P, C, t.table.as<E>(),
          res.as<E>(), nullptr);
    else
      dj_slice_kernel<E, DJ_PROBE><<<g, DJ_PNT, 0, st>>>(
          o.as<unsigned>(), starts.as<unsigned short>(), nt, t.P, C, t.table.as<E>(),
          res.as<E>(), nullptr);
    FHIP_CHECK(hipGetLastError());
  }
}

template <class K, class E>
int dj_run(const K* left, const size_t* lidx, size_t nl, const K* right, const size_t* ridx,
           size_t nr, DenseTable& t, size_t* lo, size_t* ro, size_t capacity, size_t* npairs,
           size_t* only_left, size_t* n_only, hipStream_t st) {
  if (!dj_build<K, E>(right, ridx, nr, t, st)) return DJ_DUP;
  *npairs = 0;
  if (n_only) *n_only = 0;
  if (nl == 0) return DJ_DONE;
  DBuf dest, res, starts;  // starts row P: the results in each tile
  dj_probe<K, E>(t, left, nl, dest, res, starts, st);
  const E empty = E(~E(0));
  const size_t nt = ceil_div(nl, size_t(DJ_TILE));
  FHIP_REQUIRE(nt < (size_t(1) << 31), FHIP_ERR_TOO_LARGE, "equi_join: too many probe rows");
  DBuf ws((nt + 2) * sizeof(unsigned long long));
  unsigned long long* state = ws.as<unsigned long long>();
  FHIP_CHECK(hipMemsetAsync(state, 0, (nt + 2) * sizeof(unsigned long long), st));
  {
    ProfScope prof("join_dense_emit", st);
    dj_emit_kernel<E><<<unsigned(nt), DJ_NT, 0, st>>>(
        dest.as<unsigned short>(), nl, res.as<E>(),
        starts.as<unsigned short>() + size_t(t.P) * nt, empty, t.check_empty, t.rbase, lidx, lo,
        ro, capacity, state, reinterpret_cast<unsigned*>(state + nt), state + nt + 1, only_left);
    FHIP_CHECK(hipGetLastError());
  }
  unsigned long long* h = static_cast<unsigned long long*>(pinned_scratch(0, 8));
  FHIP_CHECK(hipMemcpyAsync(h, state + nt + 1, 8, hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  *npairs = size_t(*h);
  if (n_only) *n_only = nl - size_t(*h);
  return DJ_DONE;
}

template <class K>
int dense_join_t(const K* left, const size_t* lidx, size_t nl, const K* right,
                 const size_t* ridx, size_t nr, size_t* lo, size_t* ro, size_t capacity,
                 size_t* npairs, size_t* only_left, size_t* n_only, hipStream_t st) {
  DenseTable t;
  bool wide = false;
  // 32-bit row positions within the partitioned copies
  if (nl > (size_t(1) << 32) - DJ_TILE || nr > (size_t(1) << 32) - DJ_TILE) return DJ_NA;
  if (!dj_plan<K>(right, ridx, nr, t, wide, st)) return DJ_NA;
  if (wide)
    return dj_run<K, unsigned long long>(left, lidx, nl, right, ridx, nr, t, lo, ro, capacity,
                                         npairs, only_left, n_only, st);
  return dj_run<K, unsigned>(left, lidx, nl, right, ridx, nr, t, lo, ro, capacity, npairs,
                             only_left, n_only, st);
}

template <class F>
int dj_dispatch(int ktype, F&& f) {
  int r = DJ_NA;
  dispatch_type(ktype, [&](auto z) {
    using K = decltype(z);
    if constexpr (std::is_integral<K>::value) r = f(z);
  });
  return r;
}

}  // namespace

int dense_join_pairs(int ktype, const void* left, const size_t* lidx, size_t nl,
                     const void* right, const size_t* ridx, size_t nr, size_t* lo, size_t* ro,
                     size_t capacity, size_t* npairs, size_t* only_left, size_t* n_only,
                     hipStream_t st) {
  return dj_dispatch(ktype, [&](auto z) {
    using K = decltype(z);
    return dense_join_t<K>((const K*)left, lidx, nl, (const K*)right, ridx, nr, lo, ro, capacity,
                           npairs, only_left, n_only, st);
  });
}

}  // namespace fhip
