// Stable LSD radix sort for gfx950 (one rank's share of
// core/radix_sort.hpp:197-242 and core/radix_sort.incl:7-240).
//
// Design (HBM-bound, no MFMA):
//  1. rs_hist: one read of the keys builds the 256-bin digit histogram of
//     every byte position at once (LDS counters, wave-uniform fast path).
//  2. rs_scan: exclusive scan of each histogram -> global digit bases.
//  3. rs_onesweep, once per non-trivial byte: a workgroup takes the next tile
//     (512 threads x 16 keys by default; the tile shape is tunable) in
//     arrival order, ranks its keys stably with wave64 ballot
//     matching, publishes per-digit tile counts, finds its global offsets by
//     decoupled look-back over earlier tiles, reorders the tile in LDS by
//     digit and writes keys, then payloads, in digit runs (coalesced).
// Passes whose digit is the same for every key are skipped; that covers the
// reference's "bytes of the maximum key" rule (get_max_key_size,
// radix_sort.hpp:135-153) without a separate max pass.
#include "radix_sort.hpp"

#include "fhip_device.hpp"

namespace fhip {
namespace {

constexpr int RS_NT = 256;
constexpr unsigned long long LB_FLAG_A = 1ull << 62;  // tile aggregate only
constexpr unsigned long long LB_FLAG_P = 2ull << 62;  // inclusive prefix
constexpr unsigned long long LB_VALUE = (1ull << 62) - 1;

template <int B> struct UInt;
template <> struct UInt<4> { using T = uint32_t; };
template <> struct UInt<8> { using T = uint64_t; };

template <class U>
__device__ __forceinline__ U key_xform(U k, int mode) {
  constexpr int SH = int(sizeof(U) * 8 - 1);
  const U top = U(1) << SH;
  if (mode == KM_FLIPTOP) return k ^ top;
  if (mode == KM_FLOATFLIP) return k ^ ((U(0) - (k >> SH)) | top);
  return k;
}

template <class U>
__device__ __forceinline__ unsigned key_digit(U k, int mode, int shift, int desc) {
  unsigned d = unsigned(key_xform(k, mode) >> shift) & 0xFFu;
  return desc ? 255u - d : d;
}

__device__ __forceinline__ unsigned long long lb_load(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lb_store(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- key statistics: max (init 0, as get_max_key_size) and "any < 0" ------
template <class K>
__global__ __launch_bounds__(256) void rs_stats(const K* __restrict__ keys, size_t n,
                                                unsigned long long* __restrict__ out) {
  using U = typename UInt<sizeof(K)>::T;
  unsigned long long mx = 0;
  int neg = 0;
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    K k = keys[i];
    if (k < K(0)) neg = 1;
    unsigned long long m;
    if constexpr (std::is_floating_point<K>::value) {
      U b;
      __builtin_memcpy(&b, &k, sizeof(U));
      m = b;
    } else {
      m = k > K(0) ? (unsigned long long)k : 0ull;
    }
    mx = m > mx ? m : mx;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long y = __shfl_xor(mx, o, 64);
    mx = y > mx ? y : mx;
  }
  unsigned long long anyneg = __ballot(neg);
  if (__lane_id() == 0) {
    if (mx) atomicMax(&out[0], mx);
    if (anyneg) atomicOr(&out[1], 1ull);
  }
}

// ---- all-digit histograms in one read -------------------------------------
// Every thread loads RH_ITEMS keys per step (independent loads in flight),
// then counts each byte position into LDS.  Same-address LDS atomics of one
// wave serialise, and skewed keys (a constant high byte, a heavy key) put
// most of a wave on one bin, so the lanes holding lane 0's digit add their
// count with one atomic when there are at least 8 of them.  HM picks how
// that match is found:
//   HM 2: lane 0's digit is shuffled per byte;
//   HM 4: lane 0's key is shuffled once per key and every byte compared
//         (the default: one shuffle per key instead of one per byte);
//   HM 3: no match, one atomic per lane (for keys known to be uniform).
constexpr int RH_ITEMS = 8;

template <class U, int HM>
__global__ __launch_bounds__(RS_NT) void rs_hist(const U* __restrict__ keys, size_t n,
                                                 int mode, int desc, int nbytes,
                                                 unsigned long long* __restrict__ ghist) {
  __shared__ unsigned h[sizeof(U) * 256];
  for (int i = threadIdx.x; i < int(sizeof(U)) * 256; i += RS_NT) h[i] = 0;
  __syncthreads();
  const unsigned lane = __lane_id();
  const size_t step = size_t(gridDim.x) * RS_NT * RH_ITEMS;
  for (size_t base = size_t(blockIdx.x) * RS_NT * RH_ITEMS + (threadIdx.x & ~63u); base < n;
       base += step) {
    U k[RH_ITEMS];
#pragma unroll
    for (int j = 0; j < RH_ITEMS; j++) {
      const size_t i = base + size_t(j) * RS_NT + lane;
      k[j] = i < n ? keys[i] : U(0);
    }
#pragma unroll
    for (int j = 0; j < RH_ITEMS; j++) {
      const size_t w0 = base + size_t(j) * RS_NT;  // the wave's first key of item j
      if (w0 >= n) break;                          // the same for the whole wave
      const bool valid = w0 + lane < n;
      // the key in the sort's order once; the byte loop then shifts by
      // constants (a descending sort counts 255 - digit)
      const U xk = key_xform(k[j], mode);
      const unsigned dmask = desc ? 0xFFu : 0u;
      U dx = 0;
      bool few = false;  // fewer than 8 lanes share any counted byte with lane 0
      if constexpr (HM == 4) {
        dx = xk ^ __shfl(xk, 0, 64);  // lane 0 is always valid here
        // a lane that matches lane 0 in some byte has a zero byte in dx
        // (bytes past nbytes forced non-zero); when fewer than 8 lanes do,
        // no byte reaches the 8 matches of the one-atomic case, and every
        // byte takes one atomic per lane without its ballot
        constexpr U ONES = U(~U(0)) / U(0xFF);  // 0x01 in every byte
        const U hi = nbytes >= int(sizeof(U)) ? U(0) : U(~U(0)) << (8 * nbytes);
        const U z = dx | hi;
        const bool has0 = ((z - ONES) & ~z & (ONES << 7)) != U(0);
        few = __popcll(__ballot(valid && has0)) < 8;
      }
#pragma unroll
      for (int b = 0; b < int(sizeof(U)); b++) {
        if (b >= nbytes) break;
        const unsigned d = (unsigned(xk >> (8 * b)) & 0xFFu) ^ dmask;
        if (HM == 3 || few) {
          if (valid) atomicAdd(&h[b * 256 + d], 1u);
        } else {
          bool eq0;
          if constexpr (HM == 4) eq0 = ((dx >> (8 * b)) & U(0xFF)) == U(0);
          else eq0 = d == unsigned(__shfl(int(d), 0, 64));
          const unsigned c0 = unsigned(__popcll(__ballot(valid && eq0)));
          if (c0 >= 8u) {
            if (lane == 0) atomicAdd(&h[b * 256 + d], c0);
            if (valid && !eq0) atomicAdd(&h[b * 256 + d], 1u);
          } else if (valid) {
            atomicAdd(&h[b * 256 + d], 1u);
          }
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nbytes * 256; i += RS_NT) {
    unsigned v = h[i];
    if (v) atomicAdd(&ghist[i], (unsigned long long)v);
  }
}

template <class U>
void launch_hist(int hm, int grid, const void* keys, size_t n, int mode, int desc, int nbytes,
                 unsigned long long* ghist, hipStream_t st) {
  const U* k = static_cast<const U*>(keys);
  if (hm == 3) rs_hist<U, 3><<<grid, RS_NT, 0, st>>>(k, n, mode, desc, nbytes, ghist);
  else if (hm == 2) rs_hist<U, 2><<<grid, RS_NT, 0, st>>>(k, n, mode, desc, nbytes, ghist);
  else rs_hist<U, 4><<<grid, RS_NT, 0, st>>>(k, n, mode, desc, nbytes, ghist);
  FHIP_CHECK(hipGetLastError());
}

// exclusive scan of each byte's histogram -> global digit bases
__global__ __launch_bounds__(256) void rs_scan(const unsigned long long* __restrict__ ghist,
                                               unsigned long long* __restrict__ gbase) {
  __shared__ unsigned long long tmp[4];
  const int b = blockIdx.x;
  unsigned long long v = ghist[b * 256 + threadIdx.x];
  unsigned long long e = block_excl_scan<256>(v, tmp, (unsigned long long*)nullptr);
  gbase[b * 256 + threadIdx.x] = e;
}

// ---- one onesweep pass -------------------------------------------------------
// NT threads (a multiple of 64, at least 256) with ITEMS keys each form a
// tile; the first 256 threads own one digit each for the counts, the scan
// and the look-back.
template <class U, int VB, int NT, int ITEMS>
struct PassLds {
  static constexpr int EB = (VB > int(sizeof(U))) ? VB : int(sizeof(U));
  static constexpr int WAVES = NT / 64;
  unsigned wcnt[WAVES][256];      // per-wave digit counts -> wave offsets
  unsigned tstart[256];           // tile-local start of each digit
  unsigned long long gdst[256];   // global position of tile-local index 0
  unsigned scan_tmp[WAVES];
  unsigned tile;
  alignas(16) unsigned char buf[NT * ITEMS * EB];
};

// UO, IOTA: the first pass of a narrow sort (sort_narrow) -- the payload is
// the key's row (32 bits), the caller's row ids (vsrc_, size_t) are only
// checked to be consecutive (*iota_bad set otherwise), and the keys are
// written as UO in the sort's order (flipped), so the later passes are
// plain 32-bit ones.
template <class U, int VB, int NT, int ITEMS, class UO = U, bool IOTA = false>
__global__ __launch_bounds__(NT) void rs_onesweep(
    const U* __restrict__ ksrc, UO* __restrict__ kdst, const void* __restrict__ vsrc_,
    void* __restrict__ vdst_, size_t n, int shift, int mode, int desc,
    const unsigned long long* __restrict__ gbase, unsigned long long* lookback,
    unsigned long long* lookback_next, unsigned* tile_counter,
    unsigned* tile_counter_next, unsigned* __restrict__ iota_bad = nullptr) {
  static_assert(!IOTA || VB == 4, "the narrow sort's payload is a 32-bit row");
  using V = typename UInt<(VB ? VB : 4)>::T;
  constexpr int WAVES = NT / 64;
  constexpr int TILE = NT * ITEMS;
  static_assert(NT >= 256 && NT % 64 == 0, "one thread per digit");
  __shared__ PassLds<U, VB, NT, ITEMS> s;
  const unsigned tid = threadIdx.x;
  const unsigned lane = __lane_id();
  const unsigned wave = tid >> 6;
  const bool owns_digit = tid < 256u;

  if (tid == 0) s.tile = atomicAdd(tile_counter, 1u);
  for (unsigned i = tid; i < WAVES * 256; i += NT) (&s.wcnt[0][0])[i] = 0;
  __syncthreads();
  const unsigned tile = s.tile;
  // clear this tile's look-back slots of the next pass (double buffering)
  if (owns_digit) lookback_next[size_t(tile) * 256 + tid] = 0;
  if (tile == 0 && tid == 0) *tile_counter_next = 0;

  const size_t tile_base = size_t(tile) * TILE;
  const size_t wave_base = tile_base + size_t(wave) * 64 * ITEMS;

  U keys[ITEMS];
  V vals[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    size_t idx = wave_base + size_t(j) * 64 + lane;
    keys[j] = idx < n ? ksrc[idx] : U(0);
  }
  // payloads: loaded once the look-back is done (their latency overlaps
  // the key reorder and write), which keeps the registers of the ranking
  // and look-back phases free of them

  // stable in-wave ranking: lanes with equal digits are matched with 8
  // ballots; ranks count earlier items first, then lower lanes.  Digit and
  // rank share one register (digit in the low 8 bits), and the LDS
  // positions below two to a register: the tile's rows, 32 keys and
  // payloads per lane at 1024 x 16, then fit the registers without spills.
  static_assert(TILE <= 65536 && ITEMS % 4 == 0, "16-bit tile positions, packed digits");
  unsigned dr[ITEMS];
  const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    size_t idx = wave_base + size_t(j) * 64 + lane;
    const bool valid = idx < n;
    unsigned d = key_digit(keys[j], mode, shift, desc);
    unsigned long long m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      unsigned long long bb = __ballot(bit);
      m &= bit ? bb : ~bb;
    }
    unsigned rank = 0;
    if (valid) {
      unsigned cnt = unsigned(__popcll(m));
      unsigned below = unsigned(__popcll(m & lt_mask));
      unsigned prior = s.wcnt[wave][d];
      rank = prior + below;
      if (below == cnt - 1) s.wcnt[wave][d] = prior + cnt;
    }
    dr[j] = d | (rank << 8);
  }
  __syncthreads();

  // per-digit totals and wave offsets (thread t < 256 owns digit t)
  unsigned total = 0;
  if (owns_digit) {
#pragma unroll
    for (int w = 0; w < WAVES; w++) {
      unsigned c = s.wcnt[w][tid];
      s.wcnt[w][tid] = total;
      total += c;
    }
  }
  unsigned long long* my_lb = lookback + size_t(tile) * 256 + tid;
  if (owns_digit) lb_store(my_lb, (tile == 0 ? LB_FLAG_P : LB_FLAG_A) | (unsigned long long)total);
  unsigned tstart = block_excl_scan<NT>(total, s.scan_tmp, (unsigned*)nullptr);
  if (owns_digit) {
    s.tstart[tid] = tstart;
    // decoupled look-back over earlier tiles for this digit; LB_WINDOW
    // predecessors are read at once (independent loads), then consumed in
    // order up to the first inclusive prefix or the first tile that has not
    // published yet.  Tile 0 always holds an inclusive prefix, so the walk
    // never goes below it.
    unsigned long long excl = 0;
    if (tile > 0) {
      constexpr int LB_WINDOW = 8;
      long long p = (long long)tile - 1;
      while (true) {
        unsigned long long v[LB_WINDOW];
#pragma unroll
        for (int w = 0; w < LB_WINDOW; w++)
          v[w] = (p - w >= 0) ? lb_load(lookback + size_t(p - w) * 256 + tid) : 0ull;
        int used = 0;
        bool done = false;
#pragma unroll
        for (int w = 0; w < LB_WINDOW; w++) {
          const unsigned long long flag = v[w] & ~LB_VALUE;
          if (flag == 0) break;
          excl += v[w] & LB_VALUE;
          used = w + 1;
          if (flag == LB_FLAG_P) {
            done = true;
            break;
          }
        }
        if (done) break;
        if (used == 0) __builtin_amdgcn_s_sleep(1);
        p -= used;
      }
      lb_store(my_lb, LB_FLAG_P | (excl + total));
    }
    s.gdst[tid] = gbase[tid] + excl - tstart;
  }
  __syncthreads();

  if constexpr (IOTA) {
    const size_t* ids = static_cast<const size_t*>(vsrc_);
    const size_t id0 = ids[0];
    bool bad = false;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
      size_t idx = wave_base + size_t(j) * 64 + lane;
      vals[j] = V(idx);
      if (idx < n && ids[idx] != id0 + idx) bad = true;
    }
    if (__ballot(bad) && lane == 0) atomicOr(iota_bad, 1u);
  } else if constexpr (VB != 0) {
    const V* vsrc = static_cast<const V*>(vsrc_);
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
      size_t idx = wave_base + size_t(j) * 64 + lane;
      vals[j] = idx < n ? vsrc[idx] : V(0);
    }
  }
  // reorder keys by digit in LDS
  U* kbuf = reinterpret_cast<U*>(s.buf);
  unsigned pos2[ITEMS / 2];  // tile positions, two 16-bit halves each
#pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    size_t idx = wave_base + size_t(j) * 64 + lane;
    const unsigned d = dr[j] & 0xFFu;
    const unsigned pos = s.tstart[d] + s.wcnt[wave][d] + (dr[j] >> 8);
    if (j & 1) pos2[j / 2] |= pos << 16;
    else pos2[j / 2] = pos;
    if (idx < n) kbuf[pos] = keys[j];
  }
  __syncthreads();
  const unsigned tile_n = unsigned((n - tile_base) < size_t(TILE) ? (n - tile_base) : size_t(TILE));
  unsigned dig4[ITEMS / 4];  // digits of the written keys, four to a register
#pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    unsigned i = unsigned(j) * NT + tid;
    unsigned d = 0;
    if (i < tile_n) {
      U k = kbuf[i];
      d = key_digit(k, mode, shift, desc);
      if constexpr (IOTA) kdst[s.gdst[d] + i] = UO(key_xform(k, mode));
      else kdst[s.gdst[d] + i] = k;
    }
    if (j % 4 == 0) dig4[j / 4] = d;
    else dig4[j / 4] |= d << (8 * (j % 4));
  }
  if constexpr (VB != 0) {
    V* vbuf = reinterpret_cast<V*>(s.buf);
    V* vdst = static_cast<V*>(vdst_);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
      size_t idx = wave_base + size_t(j) * 64 + lane;
      if (idx < n) vbuf[(pos2[j / 2] >> (16 * (j & 1))) & 0xFFFFu] = vals[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
      unsigned i = unsigned(j) * NT + tid;
      if (i < tile_n) vdst[s.gdst[(dig4[j / 4] >> (8 * (j % 4))) & 0xFFu] + i] = vbuf[i];
    }
  }
}

// ---- n < 64: the reference's insertion sort (radix_sort.hpp:31-99) -----------
template <class K, int VB>
__global__ void rs_insertion(K* __restrict__ k, void* __restrict__ v_, int n, int desc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  using V = typename UInt<(VB ? VB : 4)>::T;
  V* v = static_cast<V*>(v_);
  for (int i = 1; i < n; i++) {
    K tmp = k[i];
    bool ooo = desc ? (k[i - 1] < tmp) : (k[i - 1] > tmp);
    if (!ooo) continue;
    V vtmp = VB ? v[i] : V(0);
    int j = i;
    do {
      k[j] = k[j - 1];
      if (VB) v[j] = v[j - 1];
      j--;
    } while (j > 0 && (desc ? (k[j - 1] < tmp) : (k[j - 1] > tmp)));
    k[j] = tmp;
    if (VB) v[j] = vtmp;
  }
}

template <class K>
void launch_insertion(void* k, int vbytes, void* v, size_t n, bool desc, hipStream_t st) {
  if (vbytes == 0) rs_insertion<K, 0><<<1, 64, 0, st>>>((K*)k, v, int(n), desc);
  else if (vbytes == 4) rs_insertion<K, 4><<<1, 64, 0, st>>>((K*)k, v, int(n), desc);
  else rs_insertion<K, 8><<<1, 64, 0, st>>>((K*)k, v, int(n), desc);
  FHIP_CHECK(hipGetLastError());
}

template <class K>
void launch_stats(const void* keys, size_t n, unsigned long long* out, hipStream_t st) {
  int grid = int(std::min<size_t>(ceil_div(n, 256), size_t(fill_blocks(256, 4))));
  rs_stats<K><<<grid, 256, 0, st>>>((const K*)keys, n, out);
  FHIP_CHECK(hipGetLastError());
}

// Tile shapes of the onesweep pass (threads x keys per thread); the tunable
// "radix.tile" picks one by its keys per tile.
struct TileShape {
  int nt, items;
};
constexpr TileShape kTileShapes[] = {{256, 16}, {256, 32}, {512, 16}, {512, 24},
                                     {512, 12},  {1024, 8}, {1024, 16}, {768, 16}};
// measured on MI355X: 512 x 16 is the fastest key-only shape at 2^24 and
// 2^28 keys; with an 8-byte payload at 2^28 the order of 1024 x 16 and
// 512 x 16 differs from one MI355X to the next (2.12 against 2.21 ms per
// pass on one, 2.01 against 1.93 on another), so the payload default stays
// 1024 x 16 and the autotuner (radix.shape_kv) settles it per machine
constexpr int kDefaultShape = 2;
constexpr int kDefaultShapeKV = 6;
constexpr int kNumShapes = int(sizeof(kTileShapes) / sizeof(kTileShapes[0]));

template <class U, int VB, int NT, int ITEMS>
void launch_pass_t(const void* ks, void* kd, const void* vs, void* vd, size_t n, int shift,
                   int mode, int desc, const unsigned long long* gbase, unsigned long long* lb,
                   unsigned long long* lb_next, unsigned* ctr, unsigned* ctr_next,
                   unsigned tiles, hipStream_t st) {
  rs_onesweep<U, VB, NT, ITEMS><<<tiles, NT, 0, st>>>((const U*)ks, (U*)kd, vs, vd, n, shift,
                                                      mode, desc, gbase, lb, lb_next, ctr,
                                                      ctr_next);
  FHIP_CHECK(hipGetLastError());
}

using PassFn = void (*)(const void*, void*, const void*, void*, size_t, int, int, int,
                        const unsigned long long*, unsigned long long*, unsigned long long*,
                        unsigned*, unsigned*, unsigned, hipStream_t);

template <int NT, int ITEMS>
PassFn pick_pass_t(int kbytes, int vbytes) {
  if (kbytes == 4) {
    if (vbytes == 0) return launch_pass_t<uint32_t, 0, NT, ITEMS>;
    if (vbytes == 4) return launch_pass_t<uint32_t, 4, NT, ITEMS>;
    return launch_pass_t<uint32_t, 8, NT, ITEMS>;
  }
  if (vbytes == 0) return launch_pass_t<uint64_t, 0, NT, ITEMS>;
  if (vbytes == 4) return launch_pass_t<uint64_t, 4, NT, ITEMS>;
  return launch_pass_t<uint64_t, 8, NT, ITEMS>;
}

PassFn pick_pass(int shape, int kbytes, int vbytes) {
  switch (shape) {
    case 1: return pick_pass_t<256, 32>(kbytes, vbytes);
    case 2: return pick_pass_t<512, 16>(kbytes, vbytes);
    case 3: return pick_pass_t<512, 24>(kbytes, vbytes);
    case 4: return pick_pass_t<512, 12>(kbytes, vbytes);
    case 5: return pick_pass_t<1024, 8>(kbytes, vbytes);
    case 6: return pick_pass_t<1024, 16>(kbytes, vbytes);
    case 7: return pick_pass_t<768, 16>(kbytes, vbytes);
    default: return pick_pass_t<256, 16>(kbytes, vbytes);
  }
}

// the tunables "radix.shape_k" (keys only) / "radix.shape_kv" (with a
// payload): an index into kTileShapes; else the measured defaults
int chosen_shape(int vbytes) {
  const int dflt = vbytes ? kDefaultShapeKV : kDefaultShape;
  const long long idx = tune_get(vbytes ? "radix.shape_kv" : "radix.shape_k", dflt);
  return (idx >= 0 && idx < kNumShapes) ? int(idx) : dflt;
}

size_t bytes_of_max(unsigned long long mx, int kbytes) {
  // get_max_key_size, radix_sort.hpp:135-153
  unsigned long long tmp = 0xFF;
  for (int i = 1; i <= kbytes; i++) {
    if (mx <= tmp) return size_t(i);
    tmp = (tmp << 8) | 0xFF;
  }
  return size_t(kbytes);
}

// Core driver: histogram, pass selection and the pass schedule.
// Launch shape of the histogram pass: workgroups per CU (tunable
// "radix.hist_bpc") and how a wave matches lane 0's digit ("radix.hist_mode",
// rs_hist's HM: 4 by default, 2 or 3).
// Without the tunable, a workgroup counts about 16384 keys and at least 8
// workgroups per CU run (each ends with one global atomic per bin): at 2^28
// 64-bit keys 16384 workgroups took 0.73 ms against 0.96 for 2048; at 2^24,
// 2048 took 0.10 ms against 0.21 for 8192.
int hist_grid(size_t n, int* hm) {
  *hm = int(tune_get("radix.hist_mode", 4));
  const long long bpc = tune_get("radix.hist_bpc", -1);
  size_t cap;
  if (bpc > 0) {
    cap = size_t(fill_blocks(RS_NT, int(std::min(64LL, bpc))));
  } else {
    cap = std::min<size_t>(size_t(fill_blocks(RS_NT, 64)),
                           std::max<size_t>(size_t(fill_blocks(RS_NT, 8)), n / 16384));
  }
  return int(std::min<size_t>(ceil_div(n, RS_NT * RH_ITEMS), cap));
}

// known >= 0: bytes [0, known) are the passes (the caller knows the higher
// bytes are shared by every key); the host reads no histogram back
// ar: carve the temporaries from a caller's region (one temporary buffer:
// an in-place sort with an odd pass count ends with a copy back)
void run_lsd(int kbytes, int mode, int desc, int nbytes, const void* kin, void* kout,
             int vbytes, const void* vin, void* vout, size_t n, hipStream_t st,
             int known = -1, Arena* ar = nullptr) {
  if (known >= 0) nbytes = std::min(nbytes, known);
  const size_t hist_bytes = size_t(2) * 8 * 256 * sizeof(unsigned long long);
  DBuf hist;
  unsigned long long* ghist;
  if (ar) {
    ghist = static_cast<unsigned long long*>(ar->take(hist_bytes));
  } else {
    hist.reset(hist_bytes);
    ghist = hist.as<unsigned long long>();
  }
  unsigned long long* gbase = ghist + 8 * 256;
  FHIP_CHECK(hipMemsetAsync(ghist, 0, size_t(8) * 256 * sizeof(unsigned long long), st));
  int hm = 4;
  const int grid = hist_grid(n, &hm);
  if (nbytes > 0) {
    ProfScope prof("radix_hist", st);
    if (kbytes == 4) launch_hist<uint32_t>(hm, grid, kin, n, mode, desc, nbytes, ghist, st);
    else launch_hist<uint64_t>(hm, grid, kin, n, mode, desc, nbytes, ghist, st);
    rs_scan<<<nbytes, 256, 0, st>>>(ghist, gbase);
    FHIP_CHECK(hipGetLastError());
  }
  int passes[8];
  int P = 0;
  if (known >= 0) {
    for (int b = 0; b < nbytes; b++) passes[P++] = b;
  } else if (nbytes > 0) {
    unsigned long long h[8 * 256];  // on the stack: this path allocates nothing
    FHIP_CHECK(hipMemcpyAsync(h, ghist, size_t(nbytes) * 256 * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, st));
    FHIP_CHECK(hipStreamSynchronize(st));
    for (int b = 0; b < nbytes; b++) {
      unsigned long long mx = 0;
      for (int d = 0; d < 256; d++) mx = std::max(mx, h[size_t(b) * 256 + d]);
      if (mx < n) passes[P++] = b;  // a byte shared by every key is a no-op
    }
  }
  const size_t kbn = n * size_t(kbytes), vbn = n * size_t(vbytes);
  if (P == 0) {
    if (kout != kin) {
      FHIP_CHECK(hipMemcpyAsync(kout, kin, kbn, hipMemcpyDeviceToDevice, st));
      if (vbytes) FHIP_CHECK(hipMemcpyAsync(vout, vin, vbn, hipMemcpyDeviceToDevice, st));
    }
    return;
  }
  const bool inplace = (kout == kin);
  FHIP_REQUIRE(vbytes == 0 || inplace == (vout == vin), FHIP_ERR_ARG,
               "radix_sort: keys and payload must both be in place or both out of place");
  // destinations so that the last pass lands in (kout, vout) and no pass
  // reads and writes the same buffer
  DBuf t1k_own, t1v_own, t2k, t2v;
  void* t1k = nullptr;
  void* t1v = nullptr;
  if (ar) {
    t1k = ar->take(kbn);
    t1v = vbn ? ar->take(vbn) : nullptr;
  } else {
    t1k_own.reset(kbn);
    t1v_own.reset(vbn);
    t1k = t1k_own.get();
    t1v = t1v_own.get();
  }
  int dest[8];  // 0 = out, 1 = t1, 2 = t2
  bool copy_back = false;
  if (!inplace) {
    for (int i = 0; i < P; i++) dest[i] = ((P - 1 - i) % 2 == 0) ? 0 : 1;
  } else if (P == 1 || (P % 2 == 1 && ar)) {
    // one temporary: alternate t1 / out from t1, the last pass lands in t1
    for (int i = 0; i < P; i++) dest[i] = (i % 2 == 0) ? 1 : 0;
    copy_back = true;
  } else if (P % 2 == 0) {
    for (int i = 0; i < P; i++) dest[i] = (i % 2 == 0) ? 1 : 0;
  } else {
    t2k.reset(kbn);
    t2v.reset(vbn);
    dest[0] = 1;
    dest[1] = 2;
    for (int i = 2; i < P; i++) dest[i] = (i % 2 == 0) ? 0 : 1;
  }
  auto kbuf = [&](int d) -> void* { return d == 0 ? kout : d == 1 ? t1k : t2k.get(); };
  auto vbuf = [&](int d) -> void* { return d == 0 ? vout : d == 1 ? t1v : t2v.get(); };

  const int shape = chosen_shape(vbytes);
  const size_t tile_keys = size_t(kTileShapes[shape].nt) * kTileShapes[shape].items;
  const unsigned tiles = unsigned(ceil_div(n, tile_keys));
  FHIP_REQUIRE(ceil_div(n, tile_keys) < (1ull << 31), FHIP_ERR_TOO_LARGE, "radix_sort: too many keys");
  const size_t lb_words = size_t(tiles) * 256;
  const size_t lb_bytes = 2 * lb_words * sizeof(unsigned long long) + 256;
  DBuf lbuf;
  unsigned long long* lb0;
  if (ar) {
    lb0 = static_cast<unsigned long long*>(ar->take(lb_bytes));
  } else {
    lbuf.reset(lb_bytes);
    lb0 = lbuf.as<unsigned long long>();
  }
  unsigned long long* lb1 = lb0 + lb_words;
  unsigned* ctr = reinterpret_cast<unsigned*>(lb1 + lb_words);
  FHIP_CHECK(hipMemsetAsync(lb0, 0, lb_words * sizeof(unsigned long long), st));
  FHIP_CHECK(hipMemsetAsync(ctr, 0, 2 * sizeof(unsigned), st));

  PassFn fn = pick_pass(shape, kbytes, vbytes);
  const void* ks = kin;
  const void* vs = vin;
  for (int i = 0; i < P; i++) {
    void* kd = kbuf(dest[i]);
    void* vd = vbuf(dest[i]);
    unsigned long long* lb = (i & 1) ? lb1 : lb0;
    unsigned long long* lbn = (i & 1) ? lb0 : lb1;
    ProfScope prof("radix_pass", st);
    fn(ks, kd, vs, vd, n, 8 * passes[i], mode, desc, gbase + size_t(passes[i]) * 256, lb, lbn,
       ctr + (i & 1), ctr + ((i + 1) & 1), tiles, st);
    ks = kd;
    vs = vd;
  }
  if (copy_back) {
    FHIP_CHECK(hipMemcpyAsync(kout, t1k, kbn, hipMemcpyDeviceToDevice, st));
    if (vbytes) FHIP_CHECK(hipMemcpyAsync(vout, t1v, vbn, hipMemcpyDeviceToDevice, st));
  }
  // temporaries go back to the pool; later work on this stream is ordered
}

__global__ void zero_u64_kernel(unsigned long long* __restrict__ p, size_t n) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

struct AsyncLayout {
  size_t tiles, lb_words;
  size_t off_hist, off_base, off_lb0, off_lb1, off_ctr, off_t1k, off_t1v, total;
};

AsyncLayout async_layout(size_t n, int kbytes, int vbytes) {
  AsyncLayout L{};
  L.tiles = ceil_div(n, size_t(kTileShapes[kDefaultShape].nt) * kTileShapes[kDefaultShape].items);
  L.lb_words = L.tiles * 256;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    size_t r = o;
    o += (bytes + 255) & ~size_t(255);
    return r;
  };
  L.off_hist = take(8 * 256 * 8);
  L.off_base = take(8 * 256 * 8);
  L.off_lb0 = take(L.lb_words * 8);
  L.off_lb1 = take(L.lb_words * 8);
  L.off_ctr = take(256);
  L.off_t1k = take(n * size_t(kbytes));
  L.off_t1v = take(n * size_t(vbytes));
  L.total = o;
  return L;
}


// ---- the narrow sort (sort_narrow) -----------------------------------------
// All keys equal: the low 32 bits in the sort's order and each row, with the
// row ids' check.
template <class U>
__global__ void narrow_one_kernel(const U* __restrict__ keys, const size_t* __restrict__ ids,
                                  size_t n, int mode, uint32_t* __restrict__ k32,
                                  uint32_t* __restrict__ row, unsigned* __restrict__ iota_bad) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  const size_t id0 = ids[0];
  bool bad = false;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    k32[i] = uint32_t(key_xform(keys[i], mode));
    row[i] = uint32_t(i);
    if (ids[i] != id0 + i) bad = true;
  }
  if (__ballot(bad) && __lane_id() == 0) atomicOr(iota_bad, 1u);
}

using NarrowFirstFn = void (*)(const void*, void*, const void*, void*, size_t, int, int,
                               const unsigned long long*, unsigned long long*, unsigned long long*,
                               unsigned*, unsigned*, unsigned*, unsigned, hipStream_t);

template <class U, int NT, int ITEMS>
void launch_narrow_first(const void* ks, void* kd, const void* ids, void* rd, size_t n, int shift,
                         int mode, const unsigned long long* gbase, unsigned long long* lb,
                         unsigned long long* lb_next, unsigned* ctr, unsigned* ctr_next,
                         unsigned* iota_bad, unsigned tiles, hipStream_t st) {
  rs_onesweep<U, 4, NT, ITEMS, uint32_t, true><<<tiles, NT, 0, st>>>(
      (const U*)ks, (uint32_t*)kd, ids, rd, n, shift, mode, 0, gbase, lb, lb_next, ctr, ctr_next,
      iota_bad);
  FHIP_CHECK(hipGetLastError());
}

}  // namespace

size_t sort_async_workspace_bytes(size_t n, int kbytes, int vbytes) {
  return async_layout(n, kbytes, vbytes).total;
}

// Stream-ordered stable sort of unsigned keys on bytes [0, nbytes) with no
// host synchronisation, no memset and no allocation, so that it can be
// captured in a HIP graph: every byte position runs a pass (no trivial-pass
// skip).  kout must differ from kin (and vout from vin).
void sort_pairs_async(int kbytes, int nbytes, const void* kin, void* kout, int vbytes,
                      const void* vin, void* vout, size_t n, void* ws, hipStream_t st) {
  if (n == 0) return;
  FHIP_REQUIRE(kout != kin && (vbytes == 0 || vout != vin), FHIP_ERR_ARG,
               "sort_pairs_async: output must not alias the input");
  const AsyncLayout L = async_layout(n, kbytes, vbytes);
  char* w = static_cast<char*>(ws);
  auto* ghist = reinterpret_cast<unsigned long long*>(w + L.off_hist);
  auto* gbase = reinterpret_cast<unsigned long long*>(w + L.off_base);
  auto* lb0 = reinterpret_cast<unsigned long long*>(w + L.off_lb0);
  auto* lb1 = reinterpret_cast<unsigned long long*>(w + L.off_lb1);
  auto* ctr = reinterpret_cast<unsigned*>(w + L.off_ctr);
  void* t1k = w + L.off_t1k;
  void* t1v = w + L.off_t1v;
  zero_u64_kernel<<<grid_for_n(8 * 256), 256, 0, st>>>(ghist, 8 * 256);
  zero_u64_kernel<<<grid_for_n(L.lb_words), 256, 0, st>>>(lb0, L.lb_words);
  zero_u64_kernel<<<1, 64, 0, st>>>(reinterpret_cast<unsigned long long*>(ctr), 1);
  FHIP_CHECK(hipGetLastError());
  int hm = 4;
  const int grid = hist_grid(n, &hm);
  if (kbytes == 4) launch_hist<uint32_t>(hm, grid, kin, n, KM_RAW, 0, nbytes, ghist, st);
  else launch_hist<uint64_t>(hm, grid, kin, n, KM_RAW, 0, nbytes, ghist, st);
  rs_scan<<<nbytes, 256, 0, st>>>(ghist, gbase);
  FHIP_CHECK(hipGetLastError());
  PassFn fn = pick_pass(kDefaultShape, kbytes, vbytes);  // the workspace's tile count
  const void* ks = kin;
  const void* vs = vin;
  for (int i = 0; i < nbytes; i++) {
    const bool to_out = ((nbytes - 1 - i) % 2) == 0;
    void* kd = to_out ? kout : t1k;
    void* vd = to_out ? vout : t1v;
    unsigned long long* lb = (i & 1) ? lb1 : lb0;
    unsigned long long* lbn = (i & 1) ? lb0 : lb1;
    fn(ks, kd, vs, vd, n, 8 * i, KM_RAW, 0, gbase + size_t(i) * 256, lb, lbn, ctr + (i & 1),
       ctr + ((i + 1) & 1), unsigned(L.tiles), st);
    ks = kd;
    vs = vd;
  }
}

size_t sort_scratch_bytes(size_t n, int kbytes, int vbytes) {
  const size_t tk = size_t(kTileShapes[kDefaultShape].nt) * kTileShapes[kDefaultShape].items;
  const size_t tkv = size_t(kTileShapes[kDefaultShapeKV].nt) * kTileShapes[kDefaultShapeKV].items;
  // the smaller tile of either shape bounds the look-back words
  const size_t tiles = ceil_div(n, std::min(tk, tkv));
  auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
  return up(16) + up(size_t(2) * 8 * 256 * 8) + up(n * size_t(kbytes)) + up(n * size_t(vbytes)) +
         up(2 * tiles * 256 * 8 + 256);
}

// Narrow sort for the sort-merge join (see radix_sort.hpp).
bool sort_narrow(int ktype, const void* keys, const size_t* ids, size_t n, uint32_t* k32,
                 uint32_t* row, unsigned* hi, unsigned* iota_bad, hipStream_t st) {
  FHIP_REQUIRE(ktype == FHIP_I32 || ktype == FHIP_U32 || ktype == FHIP_I64 || ktype == FHIP_U64,
               FHIP_ERR_TYPE, "sort_narrow: integer keys only");
  FHIP_REQUIRE(n < (size_t(1) << 32), FHIP_ERR_TOO_LARGE, "sort_narrow: too many keys");
  *hi = 0;
  if (n == 0) return true;
  const int kbytes = type_bytes(ktype);
  const int mode = (ktype == FHIP_I32 || ktype == FHIP_I64) ? KM_FLIPTOP : KM_RAW;
  DBuf hist(size_t(2) * 8 * 256 * sizeof(unsigned long long));
  unsigned long long* ghist = hist.as<unsigned long long>();
  unsigned long long* gbase = ghist + 8 * 256;
  FHIP_CHECK(hipMemsetAsync(ghist, 0, size_t(8) * 256 * sizeof(unsigned long long), st));
  int hm = 4;
  const int grid = hist_grid(n, &hm);
  {
    ProfScope prof("radix_hist", st);
    if (kbytes == 4) launch_hist<uint32_t>(hm, grid, keys, n, mode, 0, kbytes, ghist, st);
    else launch_hist<uint64_t>(hm, grid, keys, n, mode, 0, kbytes, ghist, st);
    rs_scan<<<kbytes, 256, 0, st>>>(ghist, gbase);
    FHIP_CHECK(hipGetLastError());
  }
  unsigned long long* h =
      static_cast<unsigned long long*>(pinned_scratch(1, size_t(kbytes) * 256 * 8));
  FHIP_CHECK(hipMemcpyAsync(h, ghist, size_t(kbytes) * 256 * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  int passes[4];
  int P = 0;
  for (int b = 0; b < kbytes; b++) {
    int dmax = 0;
    unsigned long long mx = 0;
    for (int d = 0; d < 256; d++)
      if (h[size_t(b) * 256 + d] > mx) {
        mx = h[size_t(b) * 256 + d];
        dmax = d;
      }
    if (mx == n) {  // a byte shared by every key
      if (b >= 4) *hi |= unsigned(dmax) << (8 * (b - 4));
    } else if (b >= 4) {
      return false;  // the keys differ above their low 32 bits
    } else {
      passes[P++] = b;
    }
  }
  if (P == 0) {
    if (kbytes == 4)
      narrow_one_kernel<uint32_t><<<grid_for_n(n), 256, 0, st>>>(
          static_cast<const uint32_t*>(keys), ids, n, mode, k32, row, iota_bad);
    else
      narrow_one_kernel<uint64_t><<<grid_for_n(n), 256, 0, st>>>(
          static_cast<const uint64_t*>(keys), ids, n, mode, k32, row, iota_bad);
    FHIP_CHECK(hipGetLastError());
    return true;
  }
  // the last pass lands in (k32, row).  Tiles of 512 x 24 (two workgroups
  // a CU): at 2^29 28-bit keys with 32-bit rows 13.5 ms against 15.9 ms for
  // 1024 x 16 and 14.8 ms for 512 x 16 (the tunable radix.shape_narrow picks
  // 512 x 16 (2), 512 x 24 (3) or 1024 x 16 (6))
  const long long want = tune_get("radix.shape_narrow", 3);
  const int shape = (want == 2 || want == 6) ? int(want) : 3;
  const size_t tile_keys = size_t(kTileShapes[shape].nt) * kTileShapes[shape].items;
  const unsigned tiles = unsigned(ceil_div(n, tile_keys));
  DBuf tk(n * 4), tr(n * 4);
  const size_t lb_words = size_t(tiles) * 256;
  DBuf lbuf(2 * lb_words * sizeof(unsigned long long) + 256);
  unsigned long long* lb0 = lbuf.as<unsigned long long>();
  unsigned long long* lb1 = lb0 + lb_words;
  unsigned* ctr = reinterpret_cast<unsigned*>(lb1 + lb_words);
  FHIP_CHECK(hipMemsetAsync(lb0, 0, lb_words * sizeof(unsigned long long), st));
  FHIP_CHECK(hipMemsetAsync(ctr, 0, 2 * sizeof(unsigned), st));
  NarrowFirstFn first;
  if (shape == 2) first = kbytes == 4 ? launch_narrow_first<uint32_t, 512, 16> : launch_narrow_first<uint64_t, 512, 16>;
  else if (shape == 3) first = kbytes == 4 ? launch_narrow_first<uint32_t, 512, 24> : launch_narrow_first<uint64_t, 512, 24>;
  else first = kbytes == 4 ? launch_narrow_first<uint32_t, 1024, 16> : launch_narrow_first<uint64_t, 1024, 16>;
  PassFn rest = pick_pass(shape, 4, 4);
  const void* ks = keys;
  const void* vs = ids;
  for (int i = 0; i < P; i++) {
    const bool to_out = (P - 1 - i) % 2 == 0;
    void* kd = to_out ? static_cast<void*>(k32) : tk.get();
    void* vd = to_out ? static_cast<void*>(row) : tr.get();
    unsigned long long* lb = (i & 1) ? lb1 : lb0;
    unsigned long long* lbn = (i & 1) ? lb0 : lb1;
    ProfScope prof("radix_pass", st);
    if (i == 0)
      first(ks, kd, vs, vd, n, 8 * passes[i], mode, gbase + size_t(passes[i]) * 256, lb, lbn,
            ctr, ctr + 1, iota_bad, tiles, st);
    else
      rest(ks, kd, vs, vd, n, 8 * passes[i], KM_RAW, 0, gbase + size_t(passes[i]) * 256, lb, lbn,
           ctr + (i & 1), ctr + ((i + 1) & 1), tiles, st);
    ks = kd;
    vs = vd;
  }
  return true;
}

void sort_pairs(int ktype, const void* kin, void* kout, int vbytes, const void* vin,
                void* vout, size_t n, bool positive_only, bool descending, hipStream_t st,
                int low_bytes, Arena* ar) {
  const int kb = type_bytes(ktype);
  FHIP_REQUIRE(vbytes == 0 || vbytes == 4 || vbytes == 8, FHIP_ERR_ARG,
               "radix_sort: payload must be 0, 4 or 8 bytes");
  if (n == 0) return;
  FHIP_REQUIRE(vbytes == 0 || (vin && vout), FHIP_ERR_ARG, "radix_sort: null payload");
  FHIP_REQUIRE(kin && kout, FHIP_ERR_ARG, "radix_sort: null keys");
  if (n < 64) {  // SWITCH_INSERTION_THR on x86 (radix_sort.hpp:17)
    if (kout != kin) {
      FHIP_CHECK(hipMemcpyAsync(kout, kin, n * kb, hipMemcpyDeviceToDevice, st));
      if (vbytes) FHIP_CHECK(hipMemcpyAsync(vout, vin, n * vbytes, hipMemcpyDeviceToDevice, st));
    }
    switch (ktype) {
      case FHIP_I32: launch_insertion<int32_t>(kout, vbytes, vout, n, descending, st); break;
      case FHIP_U32: launch_insertion<uint32_t>(kout, vbytes, vout, n, descending, st); break;
      case FHIP_I64: launch_insertion<int64_t>(kout, vbytes, vout, n, descending, st); break;
      case FHIP_U64: launch_insertion<uint64_t>(kout, vbytes, vout, n, descending, st); break;
      case FHIP_F32: launch_insertion<float>(kout, vbytes, vout, n, descending, st); break;
      default: launch_insertion<double>(kout, vbytes, vout, n, descending, st); break;
    }
    return;
  }
  int mode = KM_RAW;
  int nbytes = kb;
  const bool need_stats =
      type_is_float(ktype) ? !positive_only : (type_is_signed_int(ktype) && positive_only);
  if (need_stats) {
    DBuf sbuf;
    unsigned long long* sdev;
    if (ar) {
      sdev = static_cast<unsigned long long*>(ar->take(2 * sizeof(unsigned long long)));
    } else {
      sbuf.reset(2 * sizeof(unsigned long long));
      sdev = sbuf.as<unsigned long long>();
    }
    FHIP_CHECK(hipMemsetAsync(sdev, 0, 2 * sizeof(unsigned long long), st));
    switch (ktype) {
      case FHIP_I32: launch_stats<int32_t>(kin, n, sdev, st); break;
      case FHIP_I64: launch_stats<int64_t>(kin, n, sdev, st); break;
      case FHIP_F32: launch_stats<float>(kin, n, sdev, st); break;
      default: launch_stats<double>(kin, n, sdev, st); break;
    }
    unsigned long long hs[2];
    FHIP_CHECK(hipMemcpyAsync(hs, sdev, sizeof(hs), hipMemcpyDeviceToHost, st));
    FHIP_CHECK(hipStreamSynchronize(st));
    if (type_is_float(ktype)) {
      mode = hs[1] ? KM_FLOATFLIP : KM_RAW;
    } else {
      mode = KM_RAW;  // positive_only: raw bits of the low bytes only
      nbytes = int(bytes_of_max(hs[0], kb));
    }
  } else if (type_is_signed_int(ktype)) {
    mode = KM_FLIPTOP;
  }
  const bool known = low_bytes >= 0 && !type_is_float(ktype) && !positive_only;
  run_lsd(kb, mode, descending ? 1 : 0, nbytes, kin, kout, vbytes, vin, vout, n, st,
          known ? std::min(low_bytes, kb) : -1, ar);
}

void sort_pairs_bits(int kbytes, const void* kin, void* kout, int vbytes, const void* vin,
                     void* vout, size_t n, int key_bits, hipStream_t st) {
  if (n == 0) return;
  int nbytes = std::max(1, std::min(kbytes, (key_bits + 7) / 8));
  run_lsd(kbytes, KM_RAW, 0, nbytes, kin, kout, vbytes, vin, vout, n, st);
}

}  // namespace fhip

// ---------------------------------------------------------------------------
// Scratch of the no-allocation entry point: a file-scope device array, one
// copy per GPU, held until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in symbol): the
// device address comes from hipGetSymbolAddress, once per device, with that
// device current.
// ---------------------------------------------------------------------------
namespace fhip {
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
}  // namespace fhip

static __device__ __align__(256) unsigned char scratch[fhip::kNoAllocScratchBytes];

namespace fhip {
namespace {
std::mutex g_scratch_mu;
std::map<int, void*>& scratch_ptrs() {
  static auto* m = new std::map<int, void*>();  // leaked: outlives static dtors
  return *m;
}
}  // namespace

// Device address of the current device's scratch array.
void* noalloc_scratch() {
  int dev = 0;
  FHIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = scratch_ptrs().find(dev);
  if (it != scratch_ptrs().end()) return it->second;
  void* ptr = nullptr;
  FHIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  FHIP_REQUIRE(ptr, FHIP_ERR_HIP, "radix_sort_noalloc: no device address for the scratch array");
  scratch_ptrs()[dev] = ptr;
  return ptr;
}
}  // namespace fhip

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int fhip_dev_radix_sort_noalloc(int ktype, void* keys, int vbytes, void* vals, size_t n,
                                int positive_only, int descending, void* stream) {
  return fhip::guarded([&] {
    const int kb = fhip::type_bytes(ktype);
    FHIP_REQUIRE(vbytes == 0 || vbytes == 4 || vbytes == 8, FHIP_ERR_ARG,
                 "radix_sort: payload must be 0, 4 or 8 bytes");
    const size_t need = fhip::sort_scratch_bytes(n, kb, vbytes);
    FHIP_REQUIRE(need <= fhip::kNoAllocScratchBytes, FHIP_ERR_TOO_LARGE,
                 "radix_sort_noalloc: " + std::to_string(n) + " keys need " +
                     std::to_string(need) + " bytes of scratch; the limit is " +
                     std::to_string(fhip::kNoAllocScratchBytes) +
                     " bytes (use fhip_dev_radix_sort for larger inputs)");
    hipStream_t st = fhip::as_stream(stream);
    fhip::Arena ar;
    ar.base = static_cast<char*>(fhip::noalloc_scratch());
    ar.cap = fhip::kNoAllocScratchBytes;
    fhip::sort_pairs(ktype, keys, keys, vbytes, vals, vals, n, positive_only != 0,
                     descending != 0, st, -1, &ar);
  });
}

int fhip_dev_radix_sort(int ktype, void* keys, int vbytes, void* vals, size_t n,
                        int positive_only, int descending, void* stream) {
  return fhip::guarded([&] {
    fhip::sort_pairs(ktype, keys, keys, vbytes, vals, vals, n, positive_only != 0,
                     descending != 0, fhip::as_stream(stream));
  });
}

int fhip_radix_sort(int ktype, void* keys, int vbytes, void* vals, size_t n,
                    int positive_only, int descending) {
  return fhip::guarded([&] {
    if (n == 0) return;
    const size_t kb = size_t(fhip::type_bytes(ktype)) * n;
    const size_t vb = size_t(vbytes) * n;
    hipStream_t st = fhip::as_stream(nullptr);
    fhip::DBuf dk(kb), dv(vb);
    fhip::copy_h2d(dk.get(), keys, kb, st);
    fhip::copy_h2d(dv.get(), vals, vb, st);
    fhip::sort_pairs(ktype, dk.get(), dk.get(), vbytes, dv.get(), dv.get(), n,
                     positive_only != 0, descending != 0, st);
    fhip::copy_d2h(keys, dk.get(), kb, st);
    fhip::copy_d2h(vals, dv.get(), vb, st);
  });
}

}  // extern "C"
