// Learned-CDF partition sort for gfx950 (no counterpart in the reference;
// its output equals radix_sort(K*, size_t) of core/radix_sort.hpp:221-242).
//
// Model: an FP32 MLP (mlp.hip) maps a normalised log-key to its CDF; it is
// trained on (key, rank) samples taken from a plain radix sort of an earlier
// input, then distilled into B1 coarse splitter keys (monotone by
// construction) and, like the leaves of a recursive model index, B2 fine
// splitters per coarse bucket taken from the sample keys inside it.
//
// Sort of 64-bit keys (key-only; equal keys are indistinguishable so the
// scatter need not be stable):
//   L1 count   : bucket of every key (binary search over the B1 splitters
//                in LDS), per-bucket counts                    read 8 B/key
//   L1 scatter : keys to their coarse bucket range (LDS ranks + one global
//                atomic per block and bucket)                  16 B/key
//   L2         : one workgroup per coarse bucket splits it into B2 fine
//                buckets                                       24 B/key
//   L3         : one workgroup per fine bucket sorts it in LDS (LSD radix
//                on key - bucket min, 8-bit digits, wave64 ballot ranks)
//                and writes the final positions                16 B/key
// Fine buckets over CAP keys (2x the expected size) are sorted afterwards by
// the plain radix sort (fallback); if too many keys need the fallback the
// whole input goes through the plain radix sort.
#include "fhip_device.hpp"
#include "radix_sort.hpp"

#include <cmath>
#include <memory>

namespace fhip {

void mlp_grad(const float* params, const float* xs, const float* ts, size_t S, float* grad,
              float* loss, hipStream_t st);
void mlp_forward(const float* params, const float* xs, size_t S, float* ys, hipStream_t st);
void mlp_adam(float* p, float* g, float* m, float* v, int step, float lr, hipStream_t st);
extern const int kMlpParams;
void prefix_sum(int type, const void* in, void* out, size_t n, hipStream_t st);

namespace {

constexpr int LS_B1 = 1024;        // coarse buckets (routed by the network)
constexpr int LS_NF = 65536;       // leaves (fine buckets, sample quantiles)
constexpr int L2_LDS_FAN = 2048;   // leaves per coarse bucket kept in LDS
constexpr int L2_NT = 1024;
constexpr int L2_ITEMS = 8;
constexpr int L2_CHUNK = L2_NT * L2_ITEMS;  // keys per second-level workgroup
constexpr int LS_NT = 512;
constexpr int L3_NT = 1024;
constexpr int L3_CAP = 8192;       // keys a fine bucket may hold (LDS)
constexpr int L3_ITEMS = L3_CAP / L3_NT;
constexpr int L3_B = 2048;         // interpolation sub-buckets per leaf
constexpr int L3_INS = 32;         // largest sub-bucket sorted by one thread
constexpr int LSD_NT = 256;
constexpr int LSD_ITEMS = L3_CAP / LSD_NT;

__device__ __forceinline__ uint64_t to_ordered(uint64_t k, int is_signed) {
  return is_signed ? (k ^ 0x8000000000000000ull) : k;
}
__device__ __forceinline__ uint64_t from_ordered(uint64_t u, int is_signed) {
  return is_signed ? (u ^ 0x8000000000000000ull) : u;
}

// number of splitters <= u, minus one (splitters[0] is 0)
__device__ __forceinline__ int bucket_of(const uint64_t* spl, int nb, uint64_t u) {
  int lo = 0, hi = nb;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (spl[mid] <= u) lo = mid + 1; else hi = mid;
  }
  return lo - 1;
}

// index of the last coarse splitter <= u (spl[0] = 0 <= u, splitters
// non-decreasing, LS_B1 a power of two): the same as bucket_of over LS_B1
// splitters, as a fixed-step search whose steps interleave across keys
__device__ __forceinline__ int coarse_of(const uint64_t* spl, uint64_t u) {
  int pos = 0;
#pragma unroll
  for (int step = LS_B1 / 2; step >= 1; step >>= 1)
    if (spl[pos + step] <= u) pos += step;
  return pos;
}

constexpr int L1C_ITEMS = 8;  // keys per thread per step of the count kernel

// ---- cell table: the coarse bucket without the full binary search ---------
// An ordered key u above `base` maps to the cell code of d = u - base (keys
// at or below base to cell 0): d itself below 2^(CT_M+1), else its bit
// length and the CT_M bits below its leading one, like a float with CT_M
// mantissa bits -- non-decreasing in u.  tab[cell] holds the coarse buckets
// of the cell's smallest and largest key (first | count << 16), so a key's
// bucket is found among those few splitters: the same index as coarse_of.
constexpr int CT_M = 6;
constexpr int CT_CELLS = (65 - CT_M) << CT_M;  // 3776

__host__ __device__ __forceinline__ unsigned cell_code(uint64_t d) {
  if (d < (uint64_t(1) << (CT_M + 1))) return unsigned(d);
  const int s = 63 - __builtin_clzll(d) - CT_M;  // >= 1
  return unsigned((uint64_t(s) << CT_M) + (d >> s));
}

__device__ __forceinline__ int coarse_by_cell(const uint64_t* spl, const unsigned* tab,
                                              uint64_t base, uint64_t u) {
  const unsigned e = tab[cell_code(u > base ? u - base : 0ull)];
  int lo = int(e & 0xFFFFu), hi = lo + int(e >> 16);
  // spl[lo] <= u; the last splitter <= u within [lo, hi]
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (spl[mid] <= u) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// bucket_of over m splitters (sp[0] <= every key) for N keys at once: the
// steps of the N searches interleave, so their LDS reads overlap
template <int N>
__device__ __forceinline__ void leaf_search(const uint64_t* sp, unsigned m, const uint64_t* u,
                                            int* pos) {
#pragma unroll
  for (int j = 0; j < N; j++) pos[j] = 0;
  unsigned top = 1;
  while (top * 2 <= m) top *= 2;
  for (unsigned st = top; st >= 1; st >>= 1) {
#pragma unroll
    for (int j = 0; j < N; j++) {
      const unsigned q = unsigned(pos[j]) + st;
      if (q < m && sp[q] <= u[j]) pos[j] = int(q);
    }
  }
}

// Leaf search through a cell table of the chunk's coarse bucket: the key
// range [sp[0], sp[m-1]) of its m leaf splitters (in LDS) is cut into at
// most L2_CELLS cells of 2^sh keys, and tab[cell] holds the leaves of the
// cell's first and last key (first | count << 16); a key at or above
// sp[m-1] is in the last leaf.  Gives leaf_search's answer (the last
// splitter <= key) with a short search among the cell's splitters.
constexpr int L2_CELLS = 1024;

struct LeafCells {
  uint64_t lo, last;  // sp[0], sp[m - 1]
  int sh, m;
};

// Fills tab (block-wide; the caller synchronises afterwards).
template <int NT>
__device__ __forceinline__ LeafCells leaf_cells_build(const uint64_t* sp, unsigned m,
                                                      unsigned* tab) {
  LeafCells lc;
  lc.lo = sp[0];
  lc.last = sp[m - 1];
  lc.m = int(m);
  const uint64_t R = lc.last - lc.lo;  // keys below sp[m-1]: offsets [0, R)
  const int bits = R ? 64 - __builtin_clzll(R - 1 ? R - 1 : 1) : 0;
  lc.sh = bits > 10 ? bits - 10 : 0;
  const unsigned ncells = R ? unsigned(((R - 1) >> lc.sh) + 1) : 0u;
  for (unsigned j = threadIdx.x; j < ncells; j += NT) {
    const uint64_t klo = lc.lo + (uint64_t(j) << lc.sh);
    const uint64_t span = (uint64_t(1) << lc.sh) - 1;
    const uint64_t khi = (R - 1 - (uint64_t(j) << lc.sh)) < span ? lc.last - 1 : klo + span;
    int p[2];
    const uint64_t q[2] = {klo, khi};
    leaf_search<2>(sp, m, q, p);
    tab[j] = unsigned(p[0]) | (unsigned(p[1] - p[0]) << 16);
  }
  return lc;
}

__device__ __forceinline__ int leaf_by_cell(const uint64_t* sp, const unsigned* tab,
                                            const LeafCells& lc, uint64_t u) {
  if (u >= lc.last) return lc.m - 1;
  if (u < lc.lo) return 0;  // below every splitter (leaf_search's answer too)
  const unsigned e = tab[unsigned((u - lc.lo) >> lc.sh)];
  int lo = int(e & 0xFFFFu), hi = lo + int(e >> 16);
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sp[mid] <= u) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(LS_NT) void l1_count_kernel(const uint64_t* __restrict__ keys,
                                                         size_t n, int is_signed,
                                                         const uint64_t* __restrict__ coarse,
                                                         const unsigned* __restrict__ ctab,
                                                         uint64_t cbase,
                                                         unsigned long long* __restrict__ hist) {
  __shared__ uint64_t spl[LS_B1];
  __shared__ unsigned h[LS_B1];
  __shared__ unsigned tab[CT_CELLS];
  for (int i = threadIdx.x; i < LS_B1; i += LS_NT) {
    spl[i] = coarse[i];
    h[i] = 0;
  }
  for (int i = threadIdx.x; i < CT_CELLS; i += LS_NT) tab[i] = ctab[i];
  __syncthreads();
  const size_t step = size_t(gridDim.x) * LS_NT * L1C_ITEMS;
  for (size_t base = size_t(blockIdx.x) * LS_NT * L1C_ITEMS; base < n; base += step) {
    uint64_t u[L1C_ITEMS];
#pragma unroll
    for (int j = 0; j < L1C_ITEMS; j++) {
      const size_t i = base + size_t(j) * LS_NT + threadIdx.x;
      u[j] = i < n ? to_ordered(keys[i], is_signed) : 0ull;
    }
    int pos[L1C_ITEMS];
#pragma unroll
    for (int j = 0; j < L1C_ITEMS; j++) pos[j] = coarse_by_cell(spl, tab, cbase, u[j]);
#pragma unroll
    for (int j = 0; j < L1C_ITEMS; j++) {
      const size_t i = base + size_t(j) * LS_NT + threadIdx.x;
      if (i < n) atomicAdd(&h[pos[j]], 1u);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < LS_B1; i += LS_NT)
    if (h[i]) atomicAdd(&hist[i], (unsigned long long)h[i]);
}

// The tile's keys are staged in LDS in bucket order before they are
// written, so every bucket's keys of the tile leave as one contiguous run
// (the key-only sort need not be stable: a key's place within its bucket's
// run is its LDS atomic's ticket).  One thread per coarse bucket for the
// tile starts (block scan) and the global reservation (one atomic).
constexpr int L1S_NT = 1024;
constexpr int L1S_ITEMS = 12;
constexpr int L1S_TILE = L1S_NT * L1S_ITEMS;
constexpr int L1S_LDS = L1S_TILE * (8 + 2);  // staged keys, their buckets
static_assert(L1S_NT == LS_B1 && L1S_TILE < 65536, "one coarse bucket per thread");

__global__ __launch_bounds__(L1S_NT) void l1_scatter_kernel(const uint64_t* __restrict__ keys,
                                                            size_t n, int is_signed,
                                                            const uint64_t* __restrict__ coarse,
                                                            const unsigned* __restrict__ ctab,
                                                            uint64_t cbase,
                                                            unsigned long long* __restrict__ cursor,
                                                            uint64_t* __restrict__ out) {
  extern __shared__ unsigned long long l1s_lds[];
  __shared__ uint64_t spl[LS_B1];
  __shared__ unsigned tab[CT_CELLS];
  __shared__ unsigned cnt[LS_B1];             // counts, then tile starts
  __shared__ unsigned long long base[LS_B1];  // global slot of tile position 0 of the bucket
  __shared__ unsigned scan_tmp[L1S_NT / 64];
  uint64_t* buf = reinterpret_cast<uint64_t*>(l1s_lds);
  unsigned short* bid = reinterpret_cast<unsigned short*>(buf + L1S_TILE);
  const unsigned tid = threadIdx.x;
  spl[tid] = coarse[tid];
  cnt[tid] = 0;
  for (int i = tid; i < CT_CELLS; i += L1S_NT) tab[i] = ctab[i];
  __syncthreads();
  const size_t tile = size_t(blockIdx.x) * L1S_TILE;
  const unsigned tn = unsigned(std::min<size_t>(L1S_TILE, n - tile));
  uint64_t k[L1S_ITEMS];
  int b[L1S_ITEMS];
  unsigned r[L1S_ITEMS];
#pragma unroll
  for (int j = 0; j < L1S_ITEMS; j++) {
    const unsigned i = unsigned(j) * L1S_NT + tid;
    k[j] = keys[tile + (i < tn ? i : 0u)];
  }
#pragma unroll
  for (int j = 0; j < L1S_ITEMS; j++) b[j] = coarse_by_cell(spl, tab, cbase, to_ordered(k[j], is_signed));
#pragma unroll
  for (int j = 0; j < L1S_ITEMS; j++) {
    const unsigned i = unsigned(j) * L1S_NT + tid;
    r[j] = i < tn ? atomicAdd(&cnt[b[j]], 1u) : 0u;
  }
  __syncthreads();
  const unsigned c = cnt[tid];
  const unsigned ts = block_excl_scan<L1S_NT>(c, scan_tmp, (unsigned*)nullptr);
  // modulo 2^64: base + (tile position) is the bucket's global slot
  base[tid] = c ? atomicAdd(&cursor[tid], (unsigned long long)c) - ts : 0ull;
  cnt[tid] = ts;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < L1S_ITEMS; j++) {
    const unsigned i = unsigned(j) * L1S_NT + tid;
    if (i < tn) {
      const unsigned pos = cnt[b[j]] + r[j];
      buf[pos] = k[j];
      bid[pos] = (unsigned short)b[j];
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < L1S_ITEMS; j++) {
    const unsigned i = unsigned(j) * L1S_NT + tid;
    if (i < tn) out[base[bid[i]] + i] = buf[i];
  }
}

// The second level works on chunks of at most L2_CHUNK keys of one coarse
// bucket c, whose leaves (fine buckets) are [kidx[c], kidx[c+1]) with
// splitters `fine`; a heavy coarse bucket is spread over many workgroups.
// Leaf counters live in LDS when the bucket has at most L2_LDS_FAN leaves,
// else the global ones are used directly.
struct L2Item {
  unsigned long long s, e;  // key range [s, e) of the chunk in the L1 output
  unsigned c, pad;
};

// One workgroup, one thread per coarse bucket: the buckets' exclusive
// offsets (off1, and the scatter's cursors), their second-level chunks in
// bucket order (items, *nitems of them) and the largest bucket, from the
// first level's counts -- on the device, so the host does not wait.
__global__ __launch_bounds__(LS_B1) void l2_items_kernel(const unsigned long long* __restrict__ hist,
                                                         unsigned long long* __restrict__ off1,
                                                         unsigned long long* __restrict__ cursor,
                                                         L2Item* __restrict__ items,
                                                         unsigned* __restrict__ nitems,
                                                         unsigned long long* __restrict__ maxc) {
  __shared__ unsigned long long tmp[LS_B1 / 64];
  __shared__ unsigned tmpc[LS_B1 / 64];
  __shared__ unsigned long long mx[LS_B1 / 64];
  const unsigned b = threadIdx.x, lane = __lane_id();
  const unsigned long long h = hist[b];
  const unsigned nch = unsigned((h + L2_CHUNK - 1) / L2_CHUNK);
  unsigned long long all;
  unsigned call;
  const unsigned long long o = block_excl_scan<LS_B1>(h, tmp, &all);
  const unsigned ci = block_excl_scan<LS_B1>(nch, tmpc, &call);
  off1[b] = o;
  cursor[b] = o;
  if (b == LS_B1 - 1) {
    off1[LS_B1] = all;
    *nitems = call;
  }
  for (unsigned q = 0; q < nch; q++) {
    const unsigned long long c = o + (unsigned long long)q * L2_CHUNK;
    items[ci + q] = L2Item{c, c + L2_CHUNK < o + h ? c + L2_CHUNK : o + h, b, 0u};
  }
  unsigned long long v = h;
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    const unsigned long long w = __shfl_xor(v, d, 64);
    v = w > v ? w : v;
  }
  if (lane == 0) mx[b >> 6] = v;
  __syncthreads();
  if (b == 0) {
    for (int w = 1; w < LS_B1 / 64; w++) v = mx[w] > v ? mx[w] : v;
    *maxc = v;
  }
}

__global__ __launch_bounds__(L2_NT) void l2_count_kernel(const L2Item* __restrict__ items,
                                                         const unsigned* __restrict__ nitems,
                                                         const uint64_t* __restrict__ in,
                                                         const unsigned* __restrict__ kidx,
                                                         int is_signed,
                                                         const uint64_t* __restrict__ fine,
                                                         unsigned long long* __restrict__ leaf_cnt) {
  __shared__ uint64_t spl[L2_LDS_FAN];
  __shared__ unsigned cnt[L2_LDS_FAN];
  __shared__ unsigned tab[L2_CELLS];
  if (blockIdx.x >= *nitems) return;  // the grid covers the largest item count
  const L2Item it = items[blockIdx.x];
  const unsigned f0 = kidx[it.c], m = kidx[it.c + 1] - f0;
  if (m == 0 || it.s == it.e) return;
  if (m == 1) {
    if (threadIdx.x == 0) atomicAdd(&leaf_cnt[f0], it.e - it.s);
    return;
  }
  const bool in_lds = m <= unsigned(L2_LDS_FAN);
  const uint64_t* sp = in_lds ? spl : fine + f0;
  if (in_lds) {
    for (unsigned j = threadIdx.x; j < m; j += L2_NT) {
      spl[j] = fine[f0 + j];
      cnt[j] = 0;
    }
  }
  const unsigned len = unsigned(it.e - it.s);
  uint64_t u[L2_ITEMS];
  int f[L2_ITEMS];
#pragma unroll
  for (int j = 0; j < L2_ITEMS; j++) {
    const unsigned i = unsigned(j) * L2_NT + threadIdx.x;
    u[j] = i < len ? to_ordered(in[it.s + i], is_signed) : ~0ull;
  }
  __syncthreads();
  if (in_lds) {
    const LeafCells lc = leaf_cells_build<L2_NT>(spl, m, tab);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < L2_ITEMS; j++) f[j] = leaf_by_cell(spl, tab, lc, u[j]);
  } else {
    leaf_search<L2_ITEMS>(sp, m, u, f);
  }
#pragma unroll
  for (int j = 0; j < L2_ITEMS; j++) {
    const unsigned i = unsigned(j) * L2_NT + threadIdx.x;
    if (i >= len) continue;
    if (in_lds) atomicAdd(&cnt[f[j]], 1u);
    else atomicAdd(&leaf_cnt[f0 + f[j]], 1ull);
  }
  __syncthreads();
  if (in_lds)
    for (unsigned j = threadIdx.x; j < m; j += L2_NT)
      if (cnt[j]) atomicAdd(&leaf_cnt[f0 + j], (unsigned long long)cnt[j]);
}

__global__ __launch_bounds__(L2_NT) void l2_scatter_kernel(const L2Item* __restrict__ items,
                                                           const unsigned* __restrict__ nitems,
                                                           const uint64_t* __restrict__ in,
                                                           const unsigned* __restrict__ kidx,
                                                           int is_signed,
                                                           const uint64_t* __restrict__ fine,
                                                           unsigned long long* __restrict__ cursor,
                                                           uint64_t* __restrict__ out) {
  __shared__ uint64_t spl[L2_LDS_FAN];
  __shared__ unsigned cnt[L2_LDS_FAN];
  __shared__ unsigned long long base[L2_LDS_FAN];
  __shared__ unsigned tab[L2_CELLS];
  if (blockIdx.x >= *nitems) return;
  const L2Item it = items[blockIdx.x];
  const unsigned f0 = kidx[it.c], m = kidx[it.c + 1] - f0;
  if (m == 0 || it.s == it.e) return;
  const unsigned len = unsigned(it.e - it.s);
  if (m == 1) {  // one leaf: the chunk moves as a block
    if (threadIdx.x == 0) base[0] = atomicAdd(&cursor[f0], (unsigned long long)len);
    __syncthreads();
    const unsigned long long b0 = base[0];
    for (unsigned i = threadIdx.x; i < len; i += L2_NT) out[b0 + i] = in[it.s + i];
    return;
  }
  const bool in_lds = m <= unsigned(L2_LDS_FAN);
  const uint64_t* sp = in_lds ? spl : fine + f0;
  if (in_lds) {
    for (unsigned j = threadIdx.x; j < m; j += L2_NT) {
      spl[j] = fine[f0 + j];
      cnt[j] = 0;
    }
  }
  uint64_t k[L2_ITEMS], u[L2_ITEMS];
  int b[L2_ITEMS];
#pragma unroll
  for (int j = 0; j < L2_ITEMS; j++) {
    const unsigned i = unsigned(j) * L2_NT + threadIdx.x;
    k[j] = i < len ? in[it.s + i] : 0ull;
    u[j] = i < len ? to_ordered(k[j], is_signed) : ~0ull;
  }
  __syncthreads();
  if (in_lds) {
    const LeafCells lc = leaf_cells_build<L2_NT>(spl, m, tab);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < L2_ITEMS; j++) b[j] = leaf_by_cell(spl, tab, lc, u[j]);
  } else {
    leaf_search<L2_ITEMS>(sp, m, u, b);
  }
#pragma unroll
  for (int j = 0; j < L2_ITEMS; j++) {
    const unsigned i = unsigned(j) * L2_NT + threadIdx.x;
    if (i >= len) b[j] = -1;
    else if (in_lds) atomicAdd(&cnt[b[j]], 1u);
  }
  if (!in_lds) {  // rare wide bucket: one global reservation per key
#pragma unroll
    for (int j = 0; j < L2_ITEMS; j++)
      if (b[j] >= 0) out[atomicAdd(&cursor[f0 + b[j]], 1ull)] = k[j];
    return;
  }
  __syncthreads();
  for (unsigned j = threadIdx.x; j < m; j += L2_NT) {
    const unsigned c = cnt[j];
    base[j] = c ? atomicAdd(&cursor[f0 + j], (unsigned long long)c) : 0;
    cnt[j] = 0;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < L2_ITEMS; j++) {
    if (b[j] < 0) continue;
    const unsigned r = atomicAdd(&cnt[b[j]], 1u);
    out[base[b[j]] + r] = k[j];
  }
}

// One workgroup per fine bucket (leaf), sorted in LDS by interpolation:
// the leaf's keys are spread over L3_B sub-buckets by
// floor((u - min) * L3_B / (max - min + 1)) -- non-decreasing in u, so the
// sub-buckets are ordered -- counted, scattered, and every sub-bucket of at
// most L3_INS keys is insertion-sorted by one thread.  Leaves inside the
// model's CDF are close to uniform, so a sub-bucket holds a few keys.  A
// parallel check over the whole leaf then proves it sorted; a leaf that
// fails (a large sub-bucket of distinct keys) or exceeds L3_CAP keys goes
// to the overflow list and is radix-sorted from `in` afterwards.  A
// single-value leaf [fine[f], fine[f] + 1) of any size is copied.
__global__ __launch_bounds__(L3_NT) void l3_kernel(const uint64_t* __restrict__ in,
                                                   const unsigned long long* __restrict__ fine_off,
                                                   const uint64_t* __restrict__ fine, int nf,
                                                   int is_signed, uint64_t* __restrict__ out,
                                                   unsigned* __restrict__ overflow_list,
                                                   unsigned* __restrict__ overflow_n) {
  __shared__ uint64_t buf[L3_CAP];
  __shared__ unsigned cur[L3_B];
  __shared__ unsigned scan_tmp[L3_NT / 64];
  __shared__ uint64_t red[2][L3_NT / 64];
  __shared__ int bad;
  const int f = blockIdx.x;
  const size_t s = fine_off[f], e = fine_off[f + 1];
  if (e - s > size_t(L3_CAP)) {
    if (f + 1 < nf && fine[f + 1] == fine[f] + 1) {
      for (size_t i = s + threadIdx.x; i < e; i += L3_NT) out[i] = in[i];
      return;
    }
    if (threadIdx.x == 0) overflow_list[atomicAdd(overflow_n, 1u)] = unsigned(f);
    return;
  }
  const unsigned m = unsigned(e - s);
  if (m == 0) return;
  const unsigned tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
  uint64_t u[L3_ITEMS];
  uint64_t mn = ~0ull, mx = 0;
#pragma unroll
  for (int j = 0; j < L3_ITEMS; j++) {
    const unsigned i = unsigned(j) * L3_NT + tid;
    u[j] = i < m ? to_ordered(in[s + i], is_signed) : 0ull;
    if (i < m) {
      mn = u[j] < mn ? u[j] : mn;
      mx = u[j] > mx ? u[j] : mx;
    }
  }
  for (unsigned i = tid; i < unsigned(L3_B); i += L3_NT) cur[i] = 0;
  if (tid == 0) bad = 0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t a = __shfl_xor(mn, o, 64), c = __shfl_xor(mx, o, 64);
    mn = a < mn ? a : mn;
    mx = c > mx ? c : mx;
  }
  if (lane == 0) {
    red[0][wave] = mn;
    red[1][wave] = mx;
  }
  __syncthreads();
  mn = red[0][0];
  mx = red[1][0];
#pragma unroll
  for (int w = 1; w < L3_NT / 64; w++) {
    mn = red[0][w] < mn ? red[0][w] : mn;
    mx = red[1][w] > mx ? red[1][w] : mx;
  }
  if (mn == mx) {  // one value
#pragma unroll
    for (int j = 0; j < L3_ITEMS; j++) {
      const unsigned i = unsigned(j) * L3_NT + tid;
      if (i < m) out[s + i] = from_ordered(u[j], is_signed);
    }
    return;
  }
  const double scale = double(L3_B) / (double(mx - mn) + 1.0);
  unsigned bk[L3_ITEMS];
#pragma unroll
  for (int j = 0; j < L3_ITEMS; j++) {
    const unsigned i = unsigned(j) * L3_NT + tid;
    bk[j] = 0;
    if (i < m) {
      const unsigned b = unsigned(double(u[j] - mn) * scale);
      bk[j] = b < unsigned(L3_B) ? b : unsigned(L3_B - 1);
      atomicAdd(&cur[bk[j]], 1u);
    }
  }
  __syncthreads();
  // exclusive scan of the sub-bucket counts (L3_B / L3_NT consecutive each)
  constexpr int PER = L3_B / L3_NT;
  unsigned c[PER], local = 0;
#pragma unroll
  for (int q = 0; q < PER; q++) {
    c[q] = cur[tid * PER + q];
    local += c[q];
  }
  unsigned run = block_excl_scan<L3_NT>(local, scan_tmp, (unsigned*)nullptr);
#pragma unroll
  for (int q = 0; q < PER; q++) {
    cur[tid * PER + q] = run;
    run += c[q];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < L3_ITEMS; j++) {
    const unsigned i = unsigned(j) * L3_NT + tid;
    if (i < m) buf[atomicAdd(&cur[bk[j]], 1u)] = u[j];
  }
  __syncthreads();
  // cur[b] is now the end of sub-bucket b (= the start of b + 1)
  for (unsigned b = tid; b < unsigned(L3_B); b += L3_NT) {
    const unsigned lo = b ? cur[b - 1] : 0u, hi = cur[b];
    if (hi - lo < 2 || hi - lo > unsigned(L3_INS)) continue;
    for (unsigned i = lo + 1; i < hi; i++) {
      const uint64_t x = buf[i];
      unsigned j = i;
      while (j > lo && buf[j - 1] > x) {
        buf[j] = buf[j - 1];
        j--;
      }
      buf[j] = x;
    }
  }
  __syncthreads();
  for (unsigned i = tid + 1; i < m; i += L3_NT)
    if (buf[i - 1] > buf[i]) bad = 1;
  __syncthreads();
  if (bad) {
    if (tid == 0) overflow_list[atomicAdd(overflow_n, 1u)] = unsigned(f);
    return;
  }
  for (unsigned i = tid; i < m; i += L3_NT) out[s + i] = from_ordered(buf[i], is_signed);
}

// The leaves that failed the interpolation check (at most L3_CAP keys):
// one workgroup per listed leaf, LSD radix sort in LDS of (key - min) with
// 8-bit digits and wave64 ballot ranking.
// The workgroups walk the list (*nleaves entries) in strides of the grid;
// every listed leaf's key range goes to rng[2q], rng[2q + 1], and a leaf of
// more than L3_CAP keys is left for the caller.
__global__ __launch_bounds__(LSD_NT) void l3_lsd_kernel(const uint64_t* __restrict__ in,
                                                       const unsigned long long* __restrict__ fine_off,
                                                       const unsigned* __restrict__ leaves,
                                                       const unsigned* __restrict__ nleaves,
                                                       unsigned long long* __restrict__ rng,
                                                       int is_signed, uint64_t* __restrict__ out) {
  __shared__ uint64_t buf[2][L3_CAP];
  __shared__ unsigned wcnt[LSD_NT / 64][256];
  __shared__ unsigned tstart[256];
  __shared__ unsigned scan_tmp[LSD_NT / 64];
  __shared__ uint64_t red[2][LSD_NT / 64];
  const unsigned nl = *nleaves;
  for (unsigned q = blockIdx.x; q < nl; q += gridDim.x) {
    const unsigned f = leaves[q];
    const size_t s = fine_off[f], e = fine_off[f + 1];
    if (threadIdx.x == 0) {
      rng[2 * size_t(q)] = s;
      rng[2 * size_t(q) + 1] = e;
    }
    const unsigned m = unsigned(e - s);
    if (m == 0 || e - s > size_t(L3_CAP)) continue;  // uniform over the workgroup
    const unsigned tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
    // load, min/max of the ordered keys
    uint64_t mn = ~0ull, mx = 0;
    for (unsigned i = tid; i < m; i += LSD_NT) {
      uint64_t u = to_ordered(in[s + i], is_signed);
      buf[0][i] = u;
      mn = u < mn ? u : mn;
      mx = u > mx ? u : mx;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      uint64_t a = __shfl_xor(mn, o, 64), c = __shfl_xor(mx, o, 64);
      mn = a < mn ? a : mn;
      mx = c > mx ? c : mx;
    }
    if (lane == 0) {
      red[0][wave] = mn;
      red[1][wave] = mx;
    }
    __syncthreads();
    mn = red[0][0];
    mx = red[1][0];
    for (int w = 1; w < LSD_NT / 64; w++) {
      mn = red[0][w] < mn ? red[0][w] : mn;
      mx = red[1][w] > mx ? red[1][w] : mx;
    }
    const uint64_t range = mx - mn;
    const int bits = range ? 64 - __clzll(range) : 0;
    const int passes = (bits + 7) / 8;
    int cur = 0;
    const unsigned long long lt_mask = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int p = 0; p < passes; p++) {
      const int sh = 8 * p;
      for (unsigned i = tid; i < (LSD_NT / 64) * 256; i += LSD_NT) (&wcnt[0][0])[i] = 0;
      __syncthreads();
      unsigned rank[LSD_ITEMS], dig[LSD_ITEMS];
#pragma unroll
      for (int j = 0; j < LSD_ITEMS; j++) {
        const unsigned idx = wave * (64 * LSD_ITEMS) + unsigned(j) * 64 + lane;
        const bool valid = idx < m;
        const unsigned d = valid ? unsigned(((buf[cur][idx] - mn) >> sh) & 0xFF) : 0u;
        unsigned long long mk = __ballot(valid);
#pragma unroll
        for (int bb = 0; bb < 8; bb++) {
          const bool bit = (d >> bb) & 1u;
          unsigned long long bl = __ballot(bit);
          mk &= bit ? bl : ~bl;
        }
        dig[j] = d;
        rank[j] = 0;
        if (valid) {
          const unsigned c = unsigned(__popcll(mk)), below = unsigned(__popcll(mk & lt_mask));
          const unsigned prior = wcnt[wave][d];
          rank[j] = prior + below;
          if (below == c - 1) wcnt[wave][d] = prior + c;
        }
      }
      __syncthreads();
      unsigned total = 0;
#pragma unroll
      for (int w = 0; w < LSD_NT / 64; w++) {
        const unsigned c = wcnt[w][tid];
        wcnt[w][tid] = total;
        total += c;
      }
      tstart[tid] = block_excl_scan<LSD_NT>(total, scan_tmp, (unsigned*)nullptr);
      __syncthreads();
#pragma unroll
      for (int j = 0; j < LSD_ITEMS; j++) {
        const unsigned idx = wave * (64 * LSD_ITEMS) + unsigned(j) * 64 + lane;
        if (idx < m) buf[cur ^ 1][tstart[dig[j]] + wcnt[wave][dig[j]] + rank[j]] = buf[cur][idx];
      }
      __syncthreads();
      cur ^= 1;
    }
    for (unsigned i = tid; i < m; i += LSD_NT) out[s + i] = from_ordered(buf[cur][i], is_signed);
    __syncthreads();  // buf and red are reused by the next listed leaf
  }
}

__global__ void to_feature_kernel(const uint64_t* __restrict__ sorted, size_t n, size_t step,
                                  size_t S, int is_signed, uint64_t umin, double inv_span,
                                  float* __restrict__ x, float* __restrict__ t) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x; k < S; k += stride) {
    const size_t i = k * step;
    const uint64_t u = to_ordered(sorted[i], is_signed);
    x[k] = float(log2(1.0 + double(u - umin)) * inv_span);
    t[k] = float(double(i) / double(n));
  }
}

// distinct keys of a sorted column: rows whose key differs from the row
// before them (plus the first), one atomic per workgroup
__global__ __launch_bounds__(256) void count_runs_kernel(const uint64_t* __restrict__ k, size_t n,
                                                         unsigned long long* __restrict__ out) {
  const size_t stride = size_t(gridDim.x) * 256;
  unsigned long long c = 0;
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride)
    c += (i == 0 || k[i] != k[i - 1]) ? 1ull : 0ull;
  c = wave_sum(c);
  __shared__ unsigned long long ws[4];
  if (__lane_id() == 0) ws[threadIdx.x / 64] = c;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, ws[0] + ws[1] + ws[2] + ws[3]);
}

// the first row of every run of equal keys of a sorted column
struct RunStartPred {
  const uint64_t* k;
  __device__ bool operator()(size_t i) const { return i == 0 || k[i] != k[i - 1]; }
};

// run j: its key and its length (the next run's start, or n, minus its own)
__global__ __launch_bounds__(256) void run_len_kernel(const uint64_t* __restrict__ k, size_t n,
                                                      const size_t* __restrict__ starts, size_t runs,
                                                      unsigned long long* __restrict__ len,
                                                      unsigned long long* __restrict__ key) {
  const size_t stride = size_t(gridDim.x) * 256;
  for (size_t j = size_t(blockIdx.x) * 256 + threadIdx.x; j < runs; j += stride) {
    const size_t a = starts[j];
    len[j] = (j + 1 < runs ? starts[j + 1] : n) - a;
    key[j] = k[a];
  }
}

}  // namespace
}  // namespace fhip

// Trained and distilled CDF model of a 64-bit key column.
struct fhip_cdf_model {
  int ktype = FHIP_I64;
  uint64_t umin = 0, umax = 0;       // ordered-key range of the training sample
  double inv_span = 1.0;             // 1 / log2(1 + umax - umin)
  std::vector<float> params;         // MLP parameters
  std::vector<double> knots_cdf;     // distilled monotone CDF at K uniform x knots
  std::vector<uint64_t> coarse;      // B1 coarse splitters (ordered keys), each a leaf splitter
  std::vector<unsigned> kidx;        // B1+1: leaves of coarse bucket c are [kidx[c], kidx[c+1])
  std::vector<uint64_t> fine;        // NF leaf splitters (sample quantiles)
  std::vector<uint64_t> sample;      // ordered sample keys (sorted)
  size_t train_n = 0;                // rows of the training input
  size_t train_distinct = 0;         // distinct keys of the training input
  // the heaviest keys of the training input (raw key bits, the longest runs
  // of the sorted column first) and their rows, for the group-by's
  // heavy-key table
  std::vector<unsigned long long> heavy_keys, heavy_counts;
  double final_loss = 0;
  // the group-by's heavy-key table placed from heavy_keys, and the last
  // cardinality estimate (n, distinct): both are pure functions of the
  // model, kept so that repeated queries skip the host work
  fhip::HotPlanCache hot_plan;
  // the learned sort's splitters, leaf indices and cell table on a device,
  // per coarse-splitter mode
  struct DevConst {
    int dev = 0, mode = 0;
    fhip::DBuf buf;
  };
  std::mutex dev_mu;
  std::vector<std::unique_ptr<DevConst>> dev_consts;
  std::mutex card_mu;
  size_t card_n = 0;
  double card_val = -1.0;
};

namespace fhip {
namespace {

constexpr int KNOTS = 4096;

uint64_t ordered_host(uint64_t k, bool sgn) { return sgn ? (k ^ 0x8000000000000000ull) : k; }

// smallest ordered key whose model CDF reaches q
uint64_t cdf_inverse(const fhip_cdf_model& m, double q) {
  const std::vector<double>& c = m.knots_cdf;
  size_t k = size_t(std::lower_bound(c.begin(), c.end(), q) - c.begin());
  double x;
  if (k == 0) x = 0.0;
  else if (k >= c.size()) x = 1.0;
  else {
    const double c0 = c[k - 1], c1 = c[k];
    const double fr = c1 > c0 ? (q - c0) / (c1 - c0) : 1.0;
    x = (double(k - 1) + fr) / double(KNOTS - 1);
  }
  const double lg = x / m.inv_span;  // log2(1 + u - umin)
  const double span = double(m.umax - m.umin);
  double d = std::exp2(lg) - 1.0;
  if (d < 0) d = 0;
  if (d > span) d = span;
  return m.umin + uint64_t(std::ceil(d));
}

// leaves: NF splitters at the sample quantiles; a value that fills two
// quantile steps gets a leaf of its own, [q, q+1), which the leaf kernel
// copies without sorting whatever its size
std::vector<uint64_t> leaf_splitters(const std::vector<uint64_t>& smp) {
  std::vector<uint64_t> f;
  f.reserve(LS_NF);
  f.push_back(0);
  auto push = [&](uint64_t w) {
    if (int(f.size()) < LS_NF && w > f.back()) f.push_back(w);
  };
  const size_t S = smp.size();
  uint64_t prev_q = S ? smp[0] : 0;
  for (size_t j = 1; j < size_t(LS_NF) && S; j++) {
    const uint64_t q = smp[std::min(S - 1, j * S / LS_NF)];
    if (q == prev_q) {
      push(q);
      if (q != ~0ull) push(q + 1);
    } else {
      push(q);
    }
    prev_q = q;
  }
  while (int(f.size()) < LS_NF) f.push_back(f.back());
  return f;
}

// Coarse splitters snapped to leaf splitters, so that every leaf belongs to
// exactly one coarse bucket.
void snap_coarse(const std::vector<uint64_t>& want, const std::vector<uint64_t>& f,
                 std::vector<uint64_t>& coarse, std::vector<unsigned>& kidx) {
  coarse.assign(LS_B1, 0);
  kidx.assign(LS_B1 + 1, 0);
  for (int c = 1; c < LS_B1; c++) {
    size_t k = size_t(std::lower_bound(f.begin(), f.end(), want[c]) - f.begin());
    k = std::min(k, size_t(LS_NF - 1));
    k = std::max<size_t>(k, kidx[c - 1]);
    kidx[c] = unsigned(k);
    coarse[c] = f[k];
  }
  kidx[LS_B1] = LS_NF;
}

void distill(fhip_cdf_model& m, const std::vector<float>& y) {
  m.knots_cdf.assign(KNOTS, 0.0);
  double run = 0.0;
  for (int k = 0; k < KNOTS; k++) {
    double v = std::min(1.0, std::max(0.0, double(y[k])));
    run = std::max(run, v);  // monotone repair
    m.knots_cdf[k] = run;
  }
  // pin the ends to the sample range
  const double lo = m.knots_cdf.front(), hi = m.knots_cdf.back();
  for (auto& v : m.knots_cdf) v = hi > lo ? (v - lo) / (hi - lo) : 0.0;
  std::vector<uint64_t> want(LS_B1, 0);
  for (int b = 1; b < LS_B1; b++)
    want[b] = std::max(cdf_inverse(m, double(b) / LS_B1), want[b - 1]);
  m.fine = leaf_splitters(m.sample);
  snap_coarse(want, m.fine, m.coarse, m.kidx);
}

}  // namespace

constexpr size_t LS_HEAVY = 4096;  // heavy keys a model keeps

// The model's heavy keys: the LS_HEAVY longest runs of the sorted training
// column (ties in key order), with their lengths.
static void learned_heavy(fhip_cdf_model* m, const void* sorted, size_t n, hipStream_t st) {
  const size_t D = m->train_distinct;
  if (D == 0 || D == n) return;  // every key distinct: none is heavy
  const uint64_t* k = static_cast<const uint64_t*>(sorted);
  DBuf sb(D * sizeof(size_t)), lb(D * 8), kb(D * 8);
  const size_t got = compact_indices(RunStartPred{k}, n, sb.as<size_t>(), st);
  FHIP_REQUIRE(got == D, FHIP_ERR_HIP, "learned_train: run count changed");
  run_len_kernel<<<grid_for_n(D), 256, 0, st>>>(k, n, sb.as<size_t>(), D,
                                                lb.as<unsigned long long>(),
                                                kb.as<unsigned long long>());
  FHIP_CHECK(hipGetLastError());
  // longest first; the stable sort keeps equal lengths in key order
  sort_pairs(FHIP_U64, lb.get(), lb.get(), 8, kb.get(), kb.get(), D, false, true, st);
  const size_t K = std::min(D, LS_HEAVY);
  m->heavy_counts.resize(K);
  m->heavy_keys.resize(K);
  FHIP_CHECK(hipMemcpyAsync(m->heavy_counts.data(), lb.get(), K * 8, hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipMemcpyAsync(m->heavy_keys.data(), kb.get(), K * 8, hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
}

fhip_cdf_model* learned_train(int ktype, const void* sorted, size_t n, size_t samples, int steps,
                              float lr, hipStream_t st) {
  FHIP_REQUIRE(ktype == FHIP_I64 || ktype == FHIP_U64, FHIP_ERR_TYPE,
               "learned sort: 64-bit integer keys only");
  FHIP_REQUIRE(n >= 2, FHIP_ERR_ARG, "learned sort: training input too small");
  const bool sgn = ktype == FHIP_I64;
  auto* m = new fhip_cdf_model();
  m->ktype = ktype;
  m->train_n = n;
  try {
    const size_t S = std::max<size_t>(2, std::min(samples, n));
    const size_t step = n / S;
    uint64_t first = 0, last = 0;
    FHIP_CHECK(hipMemcpyAsync(&first, sorted, 8, hipMemcpyDeviceToHost, st));
    FHIP_CHECK(hipMemcpyAsync(&last, static_cast<const uint64_t*>(sorted) + (n - 1), 8,
                              hipMemcpyDeviceToHost, st));
    FHIP_CHECK(hipStreamSynchronize(st));
    m->umin = ordered_host(first, sgn);
    m->umax = ordered_host(last, sgn);
    {
      DBuf cb(8);
      FHIP_CHECK(hipMemsetAsync(cb.get(), 0, 8, st));
      count_runs_kernel<<<grid_for_n(n), 256, 0, st>>>(static_cast<const uint64_t*>(sorted), n,
                                                      cb.as<unsigned long long>());
      FHIP_CHECK(hipGetLastError());
      unsigned long long d = 0;
      FHIP_CHECK(hipMemcpyAsync(&d, cb.get(), 8, hipMemcpyDeviceToHost, st));
      FHIP_CHECK(hipStreamSynchronize(st));
      m->train_distinct = size_t(d);
    }
    learned_heavy(m, sorted, n, st);
    const double span = std::log2(1.0 + double(m->umax - m->umin));
    m->inv_span = span > 0 ? 1.0 / span : 1.0;
    // the network trains on at most 2^18 of the samples (full batch)
    const size_t St = std::min<size_t>(S, size_t(1) << 18);
    const size_t step_t = n / St;
    DBuf xb(St * 4), tb(St * 4);
    to_feature_kernel<<<grid_for_n(St), 256, 0, st>>>(static_cast<const uint64_t*>(sorted), n,
                                                      step_t, St, sgn ? 1 : 0, m->umin,
                                                      m->inv_span, xb.as<float>(), tb.as<float>());
    FHIP_CHECK(hipGetLastError());
    // the sample keys themselves (for the leaves)
    std::vector<uint64_t> host_sorted_sample(S);
    copy_strided_to_host(sorted, step, S, host_sorted_sample.data(), st);
    m->sample.resize(S);
    for (size_t i = 0; i < S; i++) m->sample[i] = ordered_host(host_sorted_sample[i], sgn);
    // initialise: x spread over [0,1), ReLU units with positive slopes
    const int P = kMlpParams;
    std::vector<float> p(P, 0.f);
    uint64_t rs = 0x243F6A8885A308D3ull;
    auto rnd = [&]() {  // splitmix64 -> [-1, 1)
      rs += 0x9E3779B97F4A7C15ull;
      uint64_t z = rs;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      return float(double(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0);
    };
    for (int j = 0; j < 64; j++) {
      p[j] = 1.0f;                              // W1
      p[64 + j] = -float(j) / 64.f;             // b1: hinge at x = j/64
    }
    for (int i = 0; i < 64 * 64; i++) p[128 + i] = rnd() * 0.125f;
    for (int j = 0; j < 64; j++) p[128 + 4096 + 64 + j] = rnd() * 0.125f;
    DBuf pb(P * 4), gb(P * 4), mb(P * 4), vb(P * 4), lb(4);
    FHIP_CHECK(hipMemcpyAsync(pb.get(), p.data(), P * 4, hipMemcpyHostToDevice, st));
    FHIP_CHECK(hipMemsetAsync(gb.get(), 0, P * 4, st));
    FHIP_CHECK(hipMemsetAsync(mb.get(), 0, P * 4, st));
    FHIP_CHECK(hipMemsetAsync(vb.get(), 0, P * 4, st));
    {
      ProfScope prof("mlp_train", st);
      for (int it = 1; it <= steps; it++) {
        FHIP_CHECK(hipMemsetAsync(lb.get(), 0, 4, st));
        mlp_grad(pb.as<float>(), xb.as<float>(), tb.as<float>(), St, gb.as<float>(),
                 lb.as<float>(), st);
        mlp_adam(pb.as<float>(), gb.as<float>(), mb.as<float>(), vb.as<float>(), it, lr, st);
      }
    }
    float loss = 0;
    FHIP_CHECK(hipMemcpyAsync(&loss, lb.get(), 4, hipMemcpyDeviceToHost, st));
    // distil: forward at the knots
    std::vector<float> xk(KNOTS), yk(KNOTS);
    for (int k = 0; k < KNOTS; k++) xk[k] = float(double(k) / (KNOTS - 1));
    DBuf xkb(KNOTS * 4), ykb(KNOTS * 4);
    FHIP_CHECK(hipMemcpyAsync(xkb.get(), xk.data(), KNOTS * 4, hipMemcpyHostToDevice, st));
    mlp_forward(pb.as<float>(), xkb.as<float>(), KNOTS, ykb.as<float>(), st);
    FHIP_CHECK(hipMemcpyAsync(yk.data(), ykb.get(), KNOTS * 4, hipMemcpyDeviceToHost, st));
    m->params.resize(P);
    FHIP_CHECK(hipMemcpyAsync(m->params.data(), pb.get(), P * 4, hipMemcpyDeviceToHost, st));
    FHIP_CHECK(hipStreamSynchronize(st));
    m->final_loss = loss;
    distill(*m, yk);
  } catch (...) {
    delete m;
    throw;
  }
  return m;
}

// Sorts n keys in place; returns the number of keys that went through the
// radix fallback (n if the whole input did).  equal_width: replace the
// model's coarse splitters by equal-width ranges of the ordered key space
// (the same pipeline without the network).
// tab[cell] = first | (last - first) << 16: the coarse buckets (last
// splitter <= key, as coarse_of) of the cell's smallest and largest key
std::vector<unsigned> cell_table(const std::vector<uint64_t>& coarse, uint64_t base) {
  auto bucket = [&](uint64_t u) {
    return int(std::upper_bound(coarse.begin(), coarse.end(), u) - coarse.begin()) - 1;
  };
  std::vector<unsigned> tab(CT_CELLS);
  for (unsigned c = 0; c < unsigned(CT_CELLS); c++) {
    uint64_t lo, hi;
    if (c < (1u << (CT_M + 1))) {
      lo = hi = c;
    } else {
      const int sh = int(c >> CT_M) - 1;
      const uint64_t top = uint64_t(c) - (uint64_t(sh) << CT_M);
      lo = top << sh;
      hi = lo + ((uint64_t(1) << sh) - 1);
    }
    // keys of the cell: base + [lo, hi] (saturated), and every key at or
    // below base for cell 0
    const uint64_t klo = c == 0 ? 0ull : (lo > ~0ull - base ? ~0ull : base + lo);
    const uint64_t khi = hi > ~0ull - base ? ~0ull : base + hi;
    const int b0 = std::max(0, bucket(klo)), b1 = std::max(0, bucket(khi));
    tab[c] = unsigned(b0) | (unsigned(b1 - b0) << 16);
  }
  return tab;
}

size_t learned_sort(const fhip_cdf_model* m, void* keys, size_t n, int coarse_mode,
                    size_t* fallback_buckets, size_t* max_coarse, hipStream_t st) {
  if (fallback_buckets) *fallback_buckets = 0;
  if (max_coarse) *max_coarse = n;
  if (n < 2) return 0;
  const bool sgn = m->ktype == FHIP_I64;
  std::vector<uint64_t> coarse = m->coarse;
  std::vector<unsigned> kidx = m->kidx;
  FHIP_REQUIRE(coarse_mode >= FHIP_COARSE_NETWORK && coarse_mode <= FHIP_COARSE_QUANTILE,
               FHIP_ERR_ARG, "learned_sort: bad coarse splitter mode");
  if (coarse_mode == FHIP_COARSE_EQUAL_WIDTH) {
    // the same pipeline with equal-width coarse buckets, no network
    std::vector<uint64_t> want(LS_B1, 0);
    const double lo = double(m->umin), hi = double(m->umax);
    for (int b = 1; b < LS_B1; b++) want[b] = uint64_t(lo + (hi - lo) * double(b) / LS_B1);
    snap_coarse(want, m->fine, coarse, kidx);
  } else if (coarse_mode == FHIP_COARSE_QUANTILE) {
    // the same pipeline without the network: the coarse splitters are the
    // training sample's quantiles b / LS_B1 (snapped to leaf splitters)
    std::vector<uint64_t> want(LS_B1, 0);
    const size_t S = m->sample.size();
    for (int b = 1; b < LS_B1 && S; b++)
      want[b] = std::max(want[b - 1], m->sample[std::min(S - 1, size_t(b) * S / LS_B1)]);
    snap_coarse(want, m->fine, coarse, kidx);
  }
  const size_t nf = LS_NF;
  // the splitters, leaf indices and cell table of this mode on the device
  // (kept by the model, one copy per device and mode)
  const uint64_t cbase = m->umin;
  int dev = 0;
  FHIP_CHECK(hipGetDevice(&dev));
  const size_t cbytes = (LS_B1 + nf) * 8 + (LS_B1 + 1) * 4 + size_t(CT_CELLS) * 4;
  uint64_t* dco = nullptr;
  {
    auto* mm = const_cast<fhip_cdf_model*>(m);
    std::lock_guard<std::mutex> lock(mm->dev_mu);
    for (auto& c : mm->dev_consts)
      if (c->dev == dev && c->mode == coarse_mode) dco = c->buf.as<uint64_t>();
    if (!dco) {
      std::unique_ptr<fhip_cdf_model::DevConst> c(new fhip_cdf_model::DevConst);
      c->dev = dev;
      c->mode = coarse_mode;
      c->buf.reset(cbytes);
      std::vector<unsigned> ctab = cell_table(coarse, cbase);
      uint64_t* d = c->buf.as<uint64_t>();
      unsigned* dkk = reinterpret_cast<unsigned*>(d + LS_B1 + nf);
      FHIP_CHECK(hipMemcpyAsync(d, coarse.data(), LS_B1 * 8, hipMemcpyHostToDevice, st));
      FHIP_CHECK(hipMemcpyAsync(d + LS_B1, m->fine.data(), nf * 8, hipMemcpyHostToDevice, st));
      FHIP_CHECK(hipMemcpyAsync(dkk, kidx.data(), (LS_B1 + 1) * 4, hipMemcpyHostToDevice, st));
      FHIP_CHECK(hipMemcpyAsync(dkk + LS_B1 + 1, ctab.data(), size_t(CT_CELLS) * 4,
                                hipMemcpyHostToDevice, st));
      FHIP_CHECK(hipStreamSynchronize(st));  // the host vectors are locals
      dco = d;
      mm->dev_consts.push_back(std::move(c));
    }
  }
  uint64_t* dfi = dco + LS_B1;
  unsigned* dk = reinterpret_cast<unsigned*>(dfi + nf);
  unsigned* dct = dk + (LS_B1 + 1);
  // second-level chunks: at most one per L2_CHUNK keys plus one per bucket
  const size_t max_items = ceil_div(n, size_t(L2_CHUNK)) + LS_B1;
  DBuf hb((LS_B1 + 1) * 8 * 3 + 16), fob((nf + 1) * 8), t1(n * 8), t2(n * 8), ovb((nf + 1) * 4),
      lcb(nf * 8 * 2), itb(max_items * sizeof(L2Item)), rngb((nf + 1) * 16);
  unsigned long long* hist = hb.as<unsigned long long>();
  unsigned long long* off1 = hist + (LS_B1 + 1);
  unsigned long long* cursor = off1 + (LS_B1 + 1);
  unsigned long long* maxc = cursor + (LS_B1 + 1);
  unsigned* nitems = reinterpret_cast<unsigned*>(maxc + 1);
  unsigned long long* fine_off = fob.as<unsigned long long>();
  unsigned long long* leaf_cnt = lcb.as<unsigned long long>();
  unsigned long long* leaf_cur = leaf_cnt + nf;
  FHIP_CHECK(hipMemsetAsync(hist, 0, (LS_B1 + 1) * 8, st));
  FHIP_CHECK(hipMemsetAsync(ovb.get(), 0, 4, st));
  FHIP_CHECK(hipMemsetAsync(leaf_cnt, 0, nf * 8, st));
  FHIP_CHECK(hipMemsetAsync(fine_off, 0, 8, st));
  uint64_t* k = static_cast<uint64_t*>(keys);
  {
    ProfScope prof("learned_count", st);
    const int grid = int(std::min<size_t>(ceil_div(n, LS_NT * 8), size_t(fill_blocks(LS_NT, 4))));
    l1_count_kernel<<<grid, LS_NT, 0, st>>>(k, n, sgn, dco, dct, cbase, hist);
    FHIP_CHECK(hipGetLastError());
  }
  // exclusive offsets of the coarse buckets and the second-level chunks,
  // made on the device
  l2_items_kernel<<<1, LS_B1, 0, st>>>(hist, off1, cursor, itb.as<L2Item>(), nitems, maxc);
  FHIP_CHECK(hipGetLastError());
  {
    ProfScope prof("learned_scatter", st);
    static bool attr = false;
    if (!attr) {
      FHIP_CHECK(hipFuncSetAttribute((const void*)l1_scatter_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, L1S_LDS));
      attr = true;
    }
    const unsigned grid = unsigned(ceil_div(n, size_t(L1S_TILE)));
    l1_scatter_kernel<<<grid, L1S_NT, L1S_LDS, st>>>(k, n, sgn, dco, dct, cbase, cursor,
                                                     t1.as<uint64_t>());
    FHIP_CHECK(hipGetLastError());
  }
  {
    ProfScope prof("learned_split", st);
    const unsigned ni = unsigned(max_items);
    l2_count_kernel<<<ni, L2_NT, 0, st>>>(itb.as<L2Item>(), nitems, t1.as<uint64_t>(), dk, sgn,
                                          dfi, leaf_cnt);
    FHIP_CHECK(hipGetLastError());
    prefix_sum(FHIP_U64, leaf_cnt, fine_off + 1, nf, st);  // fine_off[nf] = n
    FHIP_CHECK(hipMemcpyAsync(leaf_cur, fine_off, nf * 8, hipMemcpyDeviceToDevice, st));
    l2_scatter_kernel<<<ni, L2_NT, 0, st>>>(itb.as<L2Item>(), nitems, t1.as<uint64_t>(), dk, sgn,
                                            dfi, leaf_cur, t2.as<uint64_t>());
    FHIP_CHECK(hipGetLastError());
  }
  {
    ProfScope prof("learned_leaf_sort", st);
    l3_kernel<<<unsigned(nf), L3_NT, 0, st>>>(t2.as<uint64_t>(), fine_off, dfi, int(nf), sgn, k,
                                              ovb.as<unsigned>() + 1, ovb.as<unsigned>());
    FHIP_CHECK(hipGetLastError());
    // the leaves that failed the check and fit in LDS: an LSD radix sort,
    // listed leaves taken in strides of the grid; every listed leaf's key
    // range is recorded for the host
    l3_lsd_kernel<<<256, LSD_NT, 0, st>>>(t2.as<uint64_t>(), fine_off, ovb.as<unsigned>() + 1,
                                          ovb.as<unsigned>(), rngb.as<unsigned long long>(), sgn, k);
    FHIP_CHECK(hipGetLastError());
  }
  // one readback: the missed leaves' count, their first kFirst ranges and the
  // largest coarse bucket
  constexpr size_t kFirst = 256;
  unsigned long long* hb_h =
      static_cast<unsigned long long*>(pinned_scratch(0, (2 + 2 * kFirst) * 8));
  FHIP_CHECK(hipMemcpyAsync(hb_h, maxc, 8, hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipMemcpyAsync(hb_h + 1, ovb.get(), 4, hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipMemcpyAsync(hb_h + 2, rngb.get(), 2 * kFirst * 8, hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  if (max_coarse) *max_coarse = size_t(hb_h[0]);
  const unsigned nov = unsigned(hb_h[1] & 0xFFFFFFFFull);
  if (fallback_buckets) *fallback_buckets = nov;
  if (nov == 0) return 0;
  std::vector<unsigned long long> rg(2 * size_t(nov));
  std::copy(hb_h + 2, hb_h + 2 + 2 * std::min<size_t>(nov, kFirst), rg.begin());
  if (nov > kFirst) {
    FHIP_CHECK(hipMemcpyAsync(rg.data() + 2 * kFirst, rngb.as<unsigned long long>() + 2 * kFirst,
                              2 * (nov - kFirst) * 8, hipMemcpyDeviceToHost, st));
    FHIP_CHECK(hipStreamSynchronize(st));
  }
  // leaves larger than LDS: the device radix sort, leaf by leaf
  size_t fb_keys = 0, big_keys = 0, nbig = 0;
  for (unsigned q = 0; q < nov; q++) {
    const size_t len = size_t(rg[2 * q + 1] - rg[2 * q]);
    fb_keys += len;
    if (len > size_t(L3_CAP)) {
      big_keys += len;
      nbig++;
    }
  }
  if (big_keys * 4 > n || nbig > 256) {
    // too much of the input missed the model: plain radix sort of all keys
    sort_pairs(m->ktype, t2.get(), k, 0, nullptr, nullptr, n, false, false, st);
    return n;
  }
  for (unsigned q = 0; q < nov; q++) {
    const size_t a = size_t(rg[2 * q]), z = size_t(rg[2 * q + 1]);
    if (z - a <= size_t(L3_CAP)) continue;
    sort_pairs(m->ktype, t2.as<uint64_t>() + a, k + a, 0, nullptr, nullptr, z - a, false, false, st);
  }
  return fb_keys;
}

// Expected number of distinct keys among n draws from the model's CDF:
// sum over knot intervals of w * (1 - exp(-n * mass / w)) with w the number
// of integer keys in the interval.
double learned_cardinality_uncached(const fhip_cdf_model* m, size_t n) {
  // a training column whose keys were all distinct (row ids, surrogate
  // keys): its draws were not independent, and n new rows of it are
  // expected to be distinct too
  if (m->train_n > 1 && m->train_distinct == m->train_n) return double(n);
  double total = 0.0;
  uint64_t prev_key = m->umin;
  double prev_c = 0.0;
  for (int k = 1; k < KNOTS; k++) {
    const double x = double(k) / (KNOTS - 1);
    const double d = std::min(std::exp2(x / m->inv_span) - 1.0, double(m->umax - m->umin));
    const uint64_t key = m->umin + uint64_t(d);
    const double c = m->knots_cdf[k];
    const double mass = std::max(0.0, c - prev_c);
    const double w = double(key - prev_key) + (k == 1 ? 1.0 : 0.0);
    if (w > 0 && mass > 0) total += w * (1.0 - std::exp(-double(n) * mass / w));
    prev_key = key;
    prev_c = c;
  }
  return std::min(total, double(n));
}

// 4095 knots of exp2 and exp per estimate: the last (n, estimate) is kept
double learned_cardinality(const fhip_cdf_model* m, size_t n) {
  auto* mm = const_cast<fhip_cdf_model*>(m);
  std::lock_guard<std::mutex> lock(mm->card_mu);
  if (mm->card_val < 0.0 || mm->card_n != n) {
    mm->card_val = learned_cardinality_uncached(m, n);
    mm->card_n = n;
  }
  return mm->card_val;
}

}  // namespace fhip

extern "C" {

int fhip_learned_train(int ktype, const void* sorted_keys, size_t n, size_t samples, int steps,
                       float lr, fhip_cdf_model** model, void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(model && sorted_keys, FHIP_ERR_ARG, "learned_train: null pointer");
    *model = fhip::learned_train(ktype, sorted_keys, n, samples, steps, lr,
                                 fhip::as_stream(stream));
  });
}

int fhip_learned_model_info(const fhip_cdf_model* m, double* final_loss, double* cdf_knots,
                            uint64_t* coarse_splitters) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(m, FHIP_ERR_ARG, "learned_model_info: null model");
    if (final_loss) *final_loss = m->final_loss;
    if (cdf_knots) std::copy(m->knots_cdf.begin(), m->knots_cdf.end(), cdf_knots);
    if (coarse_splitters) std::copy(m->coarse.begin(), m->coarse.end(), coarse_splitters);
  });
}

int fhip_learned_model_params(const fhip_cdf_model* m, float* params) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(m && params, FHIP_ERR_ARG, "learned_model_params: null pointer");
    std::copy(m->params.begin(), m->params.end(), params);
  });
}

int fhip_dev_learned_sort(const fhip_cdf_model* m, void* keys, size_t n, int coarse_mode,
                          size_t* fallback_keys, size_t* fallback_buckets,
                          size_t* max_coarse_bucket, void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(m && (n == 0 || keys), FHIP_ERR_ARG, "learned_sort: null pointer");
    size_t fb = fhip::learned_sort(m, keys, n, coarse_mode, fallback_buckets,
                                   max_coarse_bucket, fhip::as_stream(stream));
    if (fallback_keys) *fallback_keys = fb;
  });
}

int fhip_learned_cardinality(const fhip_cdf_model* m, size_t n, double* distinct) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(m && distinct, FHIP_ERR_ARG, "learned_cardinality: null pointer");
    *distinct = fhip::learned_cardinality(m, n);
  });
}

int fhip_learned_table_hint(const fhip_cdf_model* m, size_t n, size_t* table_hint) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(m && table_hint, FHIP_ERR_ARG, "learned_table_hint: null pointer");
    const double d = fhip::learned_cardinality(m, n);
    *table_hint = size_t(std::ceil(std::max(1.0, d)));
  });
}

int fhip_learned_join_strategy(const fhip_cdf_model* m, size_t n_build, int* strategy) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(m && strategy, FHIP_ERR_ARG, "learned_join_strategy: null pointer");
    // predicted unique build keys: AUTO without the sampled duplicate test
    // (the table build decides, as in the reference); else AUTO with it
    const double d = fhip::learned_cardinality(m, n_build);
    *strategy = d >= 0.999 * double(n_build) ? FHIP_JOIN_AUTO_UNIQUE : FHIP_JOIN_AUTO;
  });
}

int fhip_learned_heavy_keys(const fhip_cdf_model* m, size_t* n_keys, double* row_frac) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(m && n_keys && row_frac, FHIP_ERR_ARG, "learned_heavy_keys: null pointer");
    *n_keys = m->heavy_keys.size();
    unsigned long long rows = 0;
    for (unsigned long long c : m->heavy_counts) rows += c;
    *row_frac = m->train_n ? double(rows) / double(m->train_n) : 0.0;
  });
}

int fhip_dev_group_by_agg_model(int ktype, const void* keys, size_t n, const fhip_agg* aggs,
                                int naggs, void* keys_out, size_t capacity, size_t* ngroups,
                                const fhip_cdf_model* model, void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(model, FHIP_ERR_ARG, "group_by_agg_model: null model");
    FHIP_REQUIRE(ngroups && (naggs == 0 || aggs), FHIP_ERR_ARG, "group_by_agg: null pointer");
    FHIP_REQUIRE(n == 0 || (keys && keys_out), FHIP_ERR_ARG, "group_by_agg: null keys");
    const size_t hint = size_t(std::ceil(std::max(1.0, fhip::learned_cardinality(model, n))));
    // the heavy keys apply to a column of the model's own key type
    fhip::HeavyKeys hk;
    const bool use = model->ktype == ktype && !model->heavy_keys.empty();
    if (use) {
      hk.keys = model->heavy_keys.data();
      hk.counts = model->heavy_counts.data();
      hk.n = model->heavy_keys.size();
      hk.total = model->train_n;
      hk.cache = const_cast<fhip::HotPlanCache*>(&model->hot_plan);
    }
    fhip::group_by_agg(ktype, keys, n, aggs, naggs, keys_out, capacity, ngroups, hint,
                       fhip::as_stream(stream), use ? &hk : nullptr);
  });
}

int fhip_learned_model_destroy(fhip_cdf_model* m) {
  return fhip::guarded([&] { delete m; });
}

}  // extern "C"
