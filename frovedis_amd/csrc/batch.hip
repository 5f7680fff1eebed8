// Batched segments: thousands of independent (sort + group sum/count) jobs
// on one device in one HIP graph launch.  No counterpart in the reference,
// which runs each job as radix_sort(key, val) (core/radix_sort.hpp:197-219)
// followed by the per-group sequential fold of dfcolumn_grouped_agg.cc:
// 153-164; the outputs here are those of that per-segment loop.
//
// Segment s holds rows [seg_off[s], seg_off[s+1]) with keys in
// [0, 2^key_bits).  Two paths give the same outputs:
//   - the work queue (the default, below): small segments and key-range
//     buckets of large ones are jobs of one persistent launch, each sorted
//     and folded in LDS;
//   - the composite key (the fallback when a bucket overflows its job, or
//     "batch.path" 1): every row gets the key (s << key_bits) | key, one
//     stable LSD sort orders all segments at once (segment-major, then key,
//     ties in input order), and one thread per group folds its run.
// Either way each group's values are added one by one in row order -- the
// same additions in the same order as the reference, so the sums are
// bit-exact.  Each sequence is stream-ordered with no host
// synchronisation, memset or allocation, and is captured once per shape in
// a HIP graph that later calls replay with a single hipGraphLaunch.
#include "fhip_device.hpp"
#include "radix_sort.hpp"

#include <memory>

namespace fhip {
namespace {

__global__ void clear_kernel(unsigned long long* __restrict__ p, size_t n) {
  for (size_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// Composite keys (segment << key_bits) | key for one tile of CB_TILE
// consecutive rows per workgroup.  The segment of the tile's first row comes
// from one binary search over seg_off; the segment boundaries inside the
// tile (at most CB_MAXB of them) are kept in LDS and every row counts those
// at or below it.  A tile with more boundaries searches per row.
constexpr int CB_NT = 256;
constexpr int CB_TILE = 4096;
constexpr int CB_MAXB = 64;

__device__ __forceinline__ size_t segment_of(const size_t* seg_off, size_t nseg, size_t i) {
  // last s < nseg with seg_off[s] <= i
  size_t lo = 0, hi = nseg - 1;
  while (lo < hi) {
    size_t mid = (lo + hi + 1) >> 1;
    if (seg_off[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

template <class CT>
__global__ __launch_bounds__(CB_NT) void composite_kernel(const int64_t* __restrict__ keys,
                                                          size_t n,
                                                          const size_t* __restrict__ seg_off,
                                                          size_t nseg, int key_bits,
                                                          CT* __restrict__ comp,
                                                          int* __restrict__ err) {
  __shared__ size_t bnd[CB_MAXB];
  __shared__ size_t first_seg;
  __shared__ int nbnd;
  const size_t t0 = size_t(blockIdx.x) * CB_TILE;
  const size_t t1 = t0 + CB_TILE < n ? t0 + CB_TILE : n;
  if (threadIdx.x == 0) {
    const size_t s0 = segment_of(seg_off, nseg, t0);
    int k = 0;
    for (size_t s = s0 + 1; s < nseg && seg_off[s] < t1; s++) {
      if (k == CB_MAXB) {
        k = CB_MAXB + 1;
        break;
      }
      bnd[k++] = seg_off[s];
    }
    first_seg = s0;
    nbnd = k;
  }
  __syncthreads();
  const uint64_t lim = key_bits >= 64 ? ~0ull : (1ull << key_bits);
  const int k = nbnd;
  for (size_t i = t0 + threadIdx.x; i < t1; i += CB_NT) {
    size_t seg;
    if (k <= CB_MAXB) {
      seg = first_seg;
      for (int j = 0; j < k; j++) seg += bnd[j] <= i ? 1 : 0;
    } else {
      seg = segment_of(seg_off, nseg, i);
    }
    const int64_t key = keys[i];
    if (key < 0 || uint64_t(key) >= lim) *err = 1;
    comp[i] = CT((uint64_t(seg) << key_bits) | (uint64_t(key) & (lim - 1)));
  }
}

template <class CT>
struct RunStart {
  const CT* c;
  __device__ bool operator()(size_t i) const { return i == 0 || c[i] != c[i - 1]; }
};

// one thread per group: fold the run in order (0 + v0 + v1 + ...)
template <class CT>
__global__ void fold_kernel(const CT* __restrict__ comp, const double* __restrict__ vals,
                            const size_t* __restrict__ starts, const unsigned long long* G_dev,
                            size_t n, int key_bits, size_t capacity, int64_t* __restrict__ gkeys,
                            double* __restrict__ gsum, uint64_t* __restrict__ gcount) {
  const size_t G = size_t(*G_dev);
  const uint64_t mask = key_bits >= 64 ? ~0ull : ((1ull << key_bits) - 1);
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x; g < G; g += stride) {
    const size_t s = starts[g], e = g + 1 < G ? starts[g + 1] : n;
    double total = 0;
    for (size_t i = s; i < e; i++) total += vals[i];
    if (g < capacity) {
      gkeys[g] = int64_t(uint64_t(comp[s]) & mask);
      gsum[g] = 0.0 + total;
      gcount[g] = e - s;
    }
  }
}

// the sorted keys without their segment bits (row-parallel, coalesced)
template <class CT>
__global__ void strip_kernel(const CT* __restrict__ comp, size_t n, int key_bits,
                             int64_t* __restrict__ keys_out) {
  const uint64_t mask = key_bits >= 64 ? ~0ull : ((1ull << key_bits) - 1);
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    keys_out[i] = int64_t(uint64_t(comp[i]) & mask);
}

// group range of every segment: gseg_off[s] = first group of segment >= s
template <class CT>
__global__ void seg_groups_kernel(const CT* __restrict__ comp, const size_t* __restrict__ starts,
                                  const unsigned long long* G_dev, size_t nseg, int key_bits,
                                  size_t* __restrict__ gseg_off) {
  const size_t G = size_t(*G_dev);
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x; g <= G; g += stride) {
    const size_t sg = g < G ? size_t(uint64_t(comp[starts[g]]) >> key_bits) : nseg;
    const size_t sp = g == 0 ? 0 : size_t(uint64_t(comp[starts[g - 1]]) >> key_bits) + 1;
    for (size_t s = sp; s <= sg && s <= nseg; s++) gseg_off[s] = g;
  }
}

// ---------------------------------------------------------------------------
// Work-queue path: four launches in one stream-ordered sequence (captured
// in a HIP graph), so that bulk data only crosses kernel boundaries (within
// a kernel, rows written on one XCD are not visible to the others' L2
// without a whole-cache write-back):
//   bq_hist  one workgroup per chunk of a large segment (more rows than
//            one LDS job holds): histogram of its keys' fine bins (the top
//            BQ_FINE_BITS bits below key_bits), written per chunk;
//   bq_plan  one workgroup per large segment: sums its chunk histograms and
//            cuts the fine bins into nb buckets of about `batch.bucket_rows`
//            rows (consecutive bins, so the buckets are consecutive key
//            ranges): a bin -> bucket map and the bucket starts;
//   bq_tile  BQ_CAP rows of a large segment per workgroup, taken by ticket:
//            stable scatter by bucket into the temporary arrays (per-tile
//            wave64 ballot ranks, decoupled look-back over the segment's
//            earlier tiles), the onesweep pass of the plain radix sort with
//            the bucket as digit;
//   bq_jobs  the persistent launch: workgroups take jobs from the queue by
//            ticket, in queue order -- a small segment, or one bucket of a
//            large one: its rows sorted in LDS by key (rows packed as
//            (key - min) << 13 | row; keys spanning at most 4096 values by
//            counting, with each row placed within its key by the packed
//            rows below it; else LSD passes of at most `batch.digit_bits`
//            bits, wave64 ballot ranks: stable), written
//            out, and every group's values folded in row order by one thread
//            (the reference's sequential fold, so the sums are bit-exact);
//            each job's group count goes through a decoupled look-back over
//            the jobs, which gives every job its output offset in (segment,
//            key) order.  A job only waits on jobs with smaller tickets,
//            which running workgroups hold, so every wait ends, and every
//            workgroup leaves once the ticket passes the queue's end.
// A job whose bucket exceeds BQ_CAP rows, or whose keys span 2^19 or more,
// sets the overflow flag, and the host then runs the composite-key path.
// ---------------------------------------------------------------------------
constexpr int BQ_NT = 1024;                 // threads of the histogram, plan and tile kernels
constexpr int BQ_WAVES = BQ_NT / 64;
constexpr int BQ_TITEMS = 16;               // rows per thread of a scatter tile
constexpr int BQ_TILE_ROWS = BQ_NT * BQ_TITEMS;
// threads of a job: one 1024-thread workgroup per CU (its LDS leaves no
// room for a second).  Two co-resident 512-thread workgroups of this kernel
// gave wrong rows now and then on the MI355X; one workgroup per CU never did.
constexpr int BJ_NT = 1024;
constexpr int BJ_WAVES = BJ_NT / 64;
constexpr int BQ_ITEMS = 8;                 // rows per thread of a job
constexpr int BQ_CAP = BJ_NT * BQ_ITEMS;    // rows one LDS job holds
constexpr int BQ_FINE_BITS = 12;
constexpr int BQ_FINE = 1 << BQ_FINE_BITS;
constexpr int BQ_IDX_BITS = 13;  // row within a job (BQ_CAP = 2^13)
constexpr unsigned long long BQ_REL_LIMIT = 1ull << (32 - BQ_IDX_BITS);
constexpr int BQ_EARLY_BITS = 15;  // key span of a job whose group count is published early
static_assert((1 << BQ_EARLY_BITS) / 8 <= BQ_CAP * 4, "the early bitmap lives in the second row buffer");
static_assert(BQ_CAP == (1 << BQ_IDX_BITS), "job rows must fit the packed row field");
// key span of a job sorted by counting: one counter per key in the LSD
// passes' digit counts, four per thread
constexpr int BQ_FAST_SPAN = BJ_WAVES * 256;
static_assert(BQ_FAST_SPAN == 4 * BJ_NT && BQ_CAP < 65536, "counting sort: 4 keys per thread, 16-bit starts");
// the counting sort's largest bucket ranked by comparisons, and the
// sentinel words after the rows that its rounds of four may read
constexpr int BQ_RANK_MAX = 64;
constexpr int BQ_RANK_PAD = 128;
static_assert(BQ_RANK_MAX + 3 < BQ_RANK_PAD, "rank rounds stay within the sentinels");

enum : unsigned { BQ_HIST = 0, BQ_TILE = 1, BQ_JOB = 2 };
constexpr unsigned BQ_SMALL = 0xFFFFFFFFu;

// HIST: seg = large ordinal, idx = global chunk, rows [a, b)
// TILE: seg = large ordinal, idx = tile within the segment, rows [a, b)
// JOB:  seg = segment, idx = job number, aux = bucket (BQ_SMALL: a whole
//       small segment, rows [a, b)); for a bucket a = large ordinal
struct BqItem {
  unsigned type, seg, idx, aux;
  unsigned long long a, b;
};

struct BqSeg {
  unsigned long long start, len;
  unsigned chunk0, nchunks, tile0, ntiles, nb;
  int fshift;
};

struct BqArgs {
  const int64_t* keys;
  const double* vals;
  const BqItem* items;
  unsigned njobs;
  const BqSeg* lsegs;
  unsigned* fine_hist;           // [chunks][BQ_FINE]
  unsigned char* bmap;           // [large][BQ_FINE]: bin -> bucket
  unsigned* bstart;              // [large][257]: bucket starts within the segment
  unsigned long long* lookback;  // [tiles][256]
  unsigned long long* job_state; // [jobs]
  int64_t* tk;
  double* tv;
  int64_t* keys_out;
  double* vals_out;
  int64_t* gkeys;
  double* gsum;
  uint64_t* gcount;
  size_t* gseg_off;
  size_t capacity, nseg;
  unsigned long long key_lim;    // keys must lie in [0, key_lim)
  int digit_bits;
  int early_bits;                // jobs spanning at most this many key bits publish early
  unsigned rank_max;             // largest group the counting sort ranks by comparisons
  unsigned* ticket;              // [0] tiles, [1] jobs
  unsigned long long* scal;      // [0] key error, [1] groups, [2] overflow
};

// LDS (bytes) of the scatter tile
constexpr int BT_OFF_BUF = 0;                                 // BQ_TILE_ROWS x 8
constexpr int BT_OFF_WCNT = BT_OFF_BUF + BQ_TILE_ROWS * 8;    // BQ_WAVES x 256 x 4
constexpr int BT_OFF_TSTART = BT_OFF_WCNT + BQ_WAVES * 256 * 4;
constexpr int BT_OFF_GDST = BT_OFF_TSTART + 256 * 4;          // 256 x 8
constexpr int BT_OFF_MAP = BT_OFF_GDST + 256 * 8;             // BQ_FINE x 1
constexpr int BT_OFF_SCAN = BT_OFF_MAP + BQ_FINE;             // BQ_WAVES x 4
constexpr int BT_LDS = BT_OFF_SCAN + 64;
// LDS (bytes) of a job: two packed-row buffers, the sorted values, ranks
constexpr int BJ_OFF_PK0 = 0;                                 // BQ_CAP x 4
constexpr int BJ_OFF_PK1 = BJ_OFF_PK0 + BQ_CAP * 4;           // (BQ_CAP + BQ_RANK_PAD) x 4
constexpr int BJ_OFF_WCNT = BJ_OFF_PK1 + (BQ_CAP + BQ_RANK_PAD) * 4;  // BJ_WAVES x 256 x 4
constexpr int BJ_OFF_TSTART = BJ_OFF_WCNT + BJ_WAVES * 256 * 4;
constexpr int BJ_OFF_SCAN = BJ_OFF_TSTART + 256 * 4;          // BQ_WAVES x 4
constexpr int BJ_OFF_MISC = BJ_OFF_SCAN + 64;                 // 8 x 8
constexpr int BJ_OFF_VIN = BJ_OFF_MISC + 64;                  // BQ_CAP x 8
constexpr int BJ_LDS = BJ_OFF_VIN + BQ_CAP * 8;
static_assert(2 * BJ_LDS > 160 * 1024, "one job workgroup per CU");
// LDS (bytes) of the histogram and of the plan
constexpr int BH_LDS = BQ_FINE * 4;
constexpr int BP_LDS = 256 * 4 + 64;

// Stable ranks of the workgroup's IT x NT elements by a digit of dbits <= 8
// bits; element (wave w, item j, lane l) is position w * 64 * IT + j * 64
// + l.  Leaves the per-wave digit offsets in wcnt and returns the digit's
// total for thread t < 256 (NT >= 256).
template <int NT, int IT>
__device__ __forceinline__ unsigned bq_rank(const unsigned (&dig)[IT], const bool (&valid)[IT],
                                            int dbits, unsigned* wcnt, unsigned (&rank)[IT]) {
  constexpr int WAVES = NT / 64;
  static_assert(NT >= 256, "one thread per digit");
  const unsigned tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
  for (unsigned i = tid; i < WAVES * 256; i += NT) wcnt[i] = 0;
  __syncthreads();
  const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
#pragma unroll
  for (int j = 0; j < IT; j++) {
    unsigned long long m = __ballot(valid[j]);
    for (int b = 0; b < dbits; b++) {
      const bool bit = (dig[j] >> b) & 1u;
      const unsigned long long bb = __ballot(bit);
      m &= bit ? bb : ~bb;
    }
    rank[j] = 0;
    if (valid[j]) {
      const unsigned cnt = unsigned(__popcll(m));
      const unsigned below = unsigned(__popcll(m & lt));
      const unsigned prior = wcnt[wave * 256 + dig[j]];
      rank[j] = prior + below;
      if (below == cnt - 1) wcnt[wave * 256 + dig[j]] = prior + cnt;
    }
  }
  __syncthreads();
  unsigned total = 0;
  if (tid < 256) {
#pragma unroll
    for (int w = 0; w < WAVES; w++) {
      const unsigned c = wcnt[w * 256 + tid];
      wcnt[w * 256 + tid] = total;
      total += c;
    }
  }
  return total;
}

__device__ void bq_hist(const BqArgs& A, const BqItem& it, unsigned char* lds) {
  unsigned* h = reinterpret_cast<unsigned*>(lds);
  const BqSeg S = A.lsegs[it.seg];
  for (int i = threadIdx.x; i < BQ_FINE; i += BQ_NT) h[i] = 0;
  __syncthreads();
  int bad = 0;
  constexpr int U = 8;  // rows per thread in flight
  for (unsigned long long r0 = it.a + threadIdx.x; r0 < it.b; r0 += BQ_NT * U) {
    int64_t k[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const unsigned long long r = r0 + (unsigned long long)u * BQ_NT;
      k[u] = r < it.b ? A.keys[r] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (r0 + (unsigned long long)u * BQ_NT >= it.b) break;
      if (k[u] < 0 || (unsigned long long)k[u] >= A.key_lim) bad = 1;
      const unsigned long long f = (unsigned long long)k[u] >> S.fshift;
      atomicAdd(&h[f < BQ_FINE ? unsigned(f) : unsigned(BQ_FINE - 1)], 1u);
    }
  }
  if (__ballot(bad) && __lane_id() == 0) atomicOr(&A.scal[0], 1ull);
  __syncthreads();
  unsigned* out = A.fine_hist + size_t(it.idx) * BQ_FINE;
  for (int i = threadIdx.x; i < BQ_FINE; i += BQ_NT) out[i] = h[i];
}

__device__ void bq_plan(const BqArgs& A, unsigned L, unsigned char* lds) {
  const BqSeg S = A.lsegs[L];
  unsigned* bcnt = reinterpret_cast<unsigned*>(lds);
  unsigned* scan = reinterpret_cast<unsigned*>(lds + 256 * 4);
  const unsigned tid = threadIdx.x;
  constexpr int PER = BQ_FINE / BQ_NT;  // consecutive bins per thread
  unsigned c[PER];
  unsigned mine = 0;
#pragma unroll
  for (int e = 0; e < PER; e++) c[e] = 0;
  for (unsigned ch = 0; ch < S.nchunks; ch++) {
    const uint4 v = *reinterpret_cast<const uint4*>(A.fine_hist + size_t(S.chunk0 + ch) * BQ_FINE +
                                                     tid * PER);
    c[0] += v.x;
    c[1] += v.y;
    c[2] += v.z;
    c[3] += v.w;
  }
#pragma unroll
  for (int e = 0; e < PER; e++) mine += c[e];
  for (unsigned i = tid; i < 256; i += BQ_NT) bcnt[i] = 0;
  unsigned run = block_excl_scan<BQ_NT>(mine, scan, (unsigned*)nullptr);
  const unsigned long long target = (S.len + S.nb - 1) / S.nb;
  unsigned char* map = A.bmap + size_t(L) * BQ_FINE;
#pragma unroll
  for (int e = 0; e < PER; e++) {
    unsigned long long b = run / target;
    if (b >= S.nb) b = S.nb - 1;
    map[tid * PER + e] = (unsigned char)b;
    if (c[e]) atomicAdd(&bcnt[b], c[e]);
    run += c[e];
  }
  __syncthreads();
  const unsigned v = tid < S.nb ? bcnt[tid] : 0u;
  const unsigned ex = block_excl_scan<BQ_NT>(v, scan, (unsigned*)nullptr);
  if (tid <= 256) A.bstart[size_t(L) * 257 + tid] = ex;
}
static_assert(BQ_FINE / BQ_NT == 4, "bq_plan reads four bins per thread");

__device__ __forceinline__ unsigned long long bq_lb_load(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void bq_lb_store(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ void bq_tile(const BqArgs& A, const BqItem& it, unsigned char* lds) {
  constexpr int IT = BQ_TITEMS;
  constexpr unsigned long long FA = 1ull << 62, FP = 2ull << 62, VAL = (1ull << 62) - 1;
  const BqSeg S = A.lsegs[it.seg];
  unsigned char* map = lds + BT_OFF_MAP;
  unsigned* wcnt = reinterpret_cast<unsigned*>(lds + BT_OFF_WCNT);
  unsigned* tstart = reinterpret_cast<unsigned*>(lds + BT_OFF_TSTART);
  unsigned long long* gdst = reinterpret_cast<unsigned long long*>(lds + BT_OFF_GDST);
  unsigned* scan = reinterpret_cast<unsigned*>(lds + BT_OFF_SCAN);
  const unsigned tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
  {
    const unsigned* src = reinterpret_cast<const unsigned*>(A.bmap + size_t(it.seg) * BQ_FINE);
    unsigned* dst = reinterpret_cast<unsigned*>(map);
    for (int i = tid; i < BQ_FINE / 4; i += BQ_NT) dst[i] = src[i];
  }
  const unsigned tile_n = unsigned(it.b - it.a);
  const unsigned wave_base = wave * 64 * IT;
  // registers: the tile's keys, then its values; the digit and rank of a
  // key share one register, the LDS positions two to a register and the
  // written keys' digits four to a register (no spills at 1024 x 16)
  static_assert(BQ_TILE_ROWS <= 65536 && IT % 4 == 0, "16-bit tile positions, packed digits");
  int64_t key[IT];
  unsigned dig[IT], rank[IT];
  bool valid[IT];
#pragma unroll
  for (int j = 0; j < IT; j++) {
    const unsigned p = wave_base + unsigned(j) * 64 + lane;
    valid[j] = p < tile_n;
    key[j] = valid[j] ? A.keys[it.a + p] : 0;
  }
  __syncthreads();  // the map is in LDS
#pragma unroll
  for (int j = 0; j < IT; j++) {
    const unsigned long long f = (unsigned long long)key[j] >> S.fshift;
    dig[j] = map[f < BQ_FINE ? unsigned(f) : unsigned(BQ_FINE - 1)];
  }
  const unsigned total = bq_rank<BQ_NT, IT>(dig, valid, 8, wcnt, rank);
  unsigned dr[IT];
#pragma unroll
  for (int j = 0; j < IT; j++) dr[j] = dig[j] | (rank[j] << 8);
  const unsigned ex = block_excl_scan<BQ_NT>(tid < 256 ? total : 0u, scan, (unsigned*)nullptr);
  if (tid < 256) {
    tstart[tid] = ex;
    unsigned long long* lbs = A.lookback + size_t(S.tile0) * 256;
    unsigned long long* mine = lbs + size_t(it.idx) * 256 + tid;
    bq_lb_store(mine, (it.idx == 0 ? FP : FA) | total);
    unsigned long long excl = 0;
    if (it.idx > 0) {
      long long p = (long long)it.idx - 1;
      for (;;) {
        unsigned long long v[8];
#pragma unroll
        for (int w = 0; w < 8; w++) v[w] = (p - w >= 0) ? bq_lb_load(lbs + size_t(p - w) * 256 + tid) : 0ull;
        int used = 0;
        bool done = false;
#pragma unroll
        for (int w = 0; w < 8; w++) {
          const unsigned long long flag = v[w] & ~VAL;
          if (flag == 0) break;
          excl += v[w] & VAL;
          used = w + 1;
          if (flag == FP) {
            done = true;
            break;
          }
        }
        if (done) break;
        if (used == 0) __builtin_amdgcn_s_sleep(1);
        p -= used;
      }
      bq_lb_store(mine, FP | (excl + total));
    }
    // bucket tid's rows start at bstart[tid] within the segment (read only
    // for buckets that hold rows of this tile)
    const unsigned long long bs = total ? A.bstart[size_t(it.seg) * 257 + tid] : 0ull;
    gdst[tid] = S.start + bs + excl - ex;
  }
  // the values: loaded now, so that their latency overlaps the key reorder
  double val[IT];
#pragma unroll
  for (int j = 0; j < IT; j++) {
    const unsigned p = wave_base + unsigned(j) * 64 + lane;
    val[j] = valid[j] ? A.vals[it.a + p] : 0.0;
  }
  __syncthreads();
  int64_t* kbuf = reinterpret_cast<int64_t*>(lds + BT_OFF_BUF);
  unsigned pos2[IT / 2];
#pragma unroll
  for (int j = 0; j < IT; j++) {
    const unsigned d = dr[j] & 0xFFu;
    const unsigned pos = tstart[d] + wcnt[wave * 256 + d] + (dr[j] >> 8);
    if (j & 1) pos2[j / 2] |= pos << 16;
    else pos2[j / 2] = pos;
    if (valid[j]) kbuf[pos] = key[j];
  }
  __syncthreads();
  unsigned dig4[IT / 4];
#pragma unroll
  for (int j = 0; j < IT; j++) {
    const unsigned i = unsigned(j) * BQ_NT + tid;
    unsigned d = 0;
    if (i < tile_n) {
      const int64_t k = kbuf[i];
      const unsigned long long f = (unsigned long long)k >> S.fshift;
      d = map[f < BQ_FINE ? unsigned(f) : unsigned(BQ_FINE - 1)];
      A.tk[gdst[d] + i] = k;
    }
    if (j % 4 == 0) dig4[j / 4] = d;
    else dig4[j / 4] |= d << (8 * (j % 4));
  }
  __syncthreads();
  double* vbuf = reinterpret_cast<double*>(lds + BT_OFF_BUF);
#pragma unroll
  for (int j = 0; j < IT; j++)
    if (valid[j]) vbuf[(pos2[j / 2] >> (16 * (j & 1))) & 0xFFFFu] = val[j];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < IT; j++) {
    const unsigned i = unsigned(j) * BQ_NT + tid;
    if (i < tile_n) A.tv[gdst[(dig4[j / 4] >> (8 * (j % 4))) & 0xFFu] + i] = vbuf[i];
  }
}

// The second half of compact_lookback for a job that has already published
// its count: sums earlier jobs' counts (64 at a time) up to the first
// inclusive prefix, then publishes its own; all 64 lanes of one wave.
__device__ __forceinline__ unsigned long long bq_job_resolve(unsigned long long* state,
                                                             unsigned idx, unsigned long long run) {
  const unsigned lane = __lane_id();
  if (idx == 0) return 0;
  unsigned long long excl = 0;
  long long p = (long long)idx - 1;  // lane l looks at job p - l
  while (true) {
    const unsigned long long v = (p - (long long)lane >= 0) ? bq_lb_load(state + (p - lane)) : CP_FLAG_P;
    const unsigned long long flag = v & ~CP_VALUE;
    const unsigned long long not_ready = __ballot(flag == 0);
    const unsigned long long prefix = __ballot(flag == CP_FLAG_P);
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
  if (lane == 0) bq_lb_store(state + idx, CP_FLAG_P | (excl + run));
  return excl;
}

__device__ void bq_job(const BqArgs& A, const BqItem& it, unsigned char* lds) {
  constexpr int IT = BQ_ITEMS;
  const unsigned tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
  // the packed-row buffer c (0 or 1): an offset from the LDS base, so the
  // accesses stay LDS instructions (a pointer picked from an array of
  // pointers would become a generic one, and FLAT accesses to LDS are not
  // ordered by the barriers' LDS waits)
  auto pk = [lds](int c) {
    return reinterpret_cast<unsigned*>(lds + BJ_OFF_PK0 + (c ? BQ_CAP * 4 : 0));
  };
  // the job's values in input order, loaded with the keys (coalesced); the
  // sorted rows read them from here (by their row field) instead of
  // gathering from memory, and so do the folds
  double* vin = reinterpret_cast<double*>(lds + BJ_OFF_VIN);
  unsigned* wcnt = reinterpret_cast<unsigned*>(lds + BJ_OFF_WCNT);
  unsigned* tstart = reinterpret_cast<unsigned*>(lds + BJ_OFF_TSTART);
  unsigned* scan = reinterpret_cast<unsigned*>(lds + BJ_OFF_SCAN);
  unsigned long long* misc = reinterpret_cast<unsigned long long*>(lds + BJ_OFF_MISC);
  const bool small = it.aux == BQ_SMALL;
  unsigned long long a, m;
  const int64_t* ksrc;
  const double* vsrc;
  if (small) {
    a = it.a;
    m = it.b - it.a;
    ksrc = A.keys + a;
    vsrc = A.vals + a;
  } else {
    const BqSeg S = A.lsegs[it.a];
    const unsigned* bs = A.bstart + size_t(it.a) * 257;
    a = S.start + bs[it.aux];
    m = bs[it.aux + 1] - bs[it.aux];
    ksrc = A.tk + a;
    vsrc = A.tv + a;
  }
  if (tid == 0) {
    misc[0] = ~0ull;  // min
    misc[1] = 0;      // max
    misc[3] = 0;      // a bucket of the counting sort larger than rank_max
  }
  for (unsigned i = tid; i < BQ_FAST_SPAN; i += BJ_NT) wcnt[i] = 0;  // counting-sort counts
  __syncthreads();
  const bool fits = m <= BQ_CAP;
  const unsigned wave_base = wave * 64 * IT;
  unsigned long long kv[IT];
  bool valid[IT];
  int bad = 0;
  unsigned long long lo = ~0ull, hi = 0;
  double xv[IT];
#pragma unroll
  for (int j = 0; j < IT; j++) {
    const unsigned p = wave_base + unsigned(j) * 64 + lane;
    valid[j] = fits && p < m;
    xv[j] = valid[j] ? vsrc[p] : 0.0;
  }
#pragma unroll
  for (int j = 0; j < IT; j++) {
    const unsigned p = wave_base + unsigned(j) * 64 + lane;
    const int64_t k = valid[j] ? ksrc[p] : 0;
    if (valid[j] && (k < 0 || (unsigned long long)k >= A.key_lim)) bad = 1;
    kv[j] = (unsigned long long)k;
    if (valid[j]) {
      lo = kv[j] < lo ? kv[j] : lo;
      hi = kv[j] > hi ? kv[j] : hi;
    }
  }
#pragma unroll
  for (int j = 0; j < IT; j++)
    if (valid[j]) vin[wave_base + unsigned(j) * 64 + lane] = xv[j];
  if (small && __ballot(bad) && lane == 0) atomicOr(&A.scal[0], 1ull);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long l2 = __shfl_xor(lo, o, 64), h2 = __shfl_xor(hi, o, 64);
    lo = l2 < lo ? l2 : lo;
    hi = h2 > hi ? h2 : hi;
  }
  if (lane == 0) {
    atomicMin(&misc[0], lo);
    atomicMax(&misc[1], hi);
  }
  __syncthreads();
  const unsigned long long kmin = misc[0];
  const unsigned long long range = m ? misc[1] - kmin : 0;
  const bool ok = fits && range < BQ_REL_LIMIT;
  if (!ok && tid == 0) atomicOr(&A.scal[2], 1ull);
  const int kbits = range ? 64 - __builtin_clzll(range) : 0;
  // a key span of at most BQ_FAST_SPAN values: a counting sort (below),
  // which gives the group count and starts from its counts
  const bool fast = ok && range < BQ_FAST_SPAN;
  unsigned pv[IT];
#pragma unroll
  for (int j = 0; j < IT; j++) {
    const unsigned p = wave_base + unsigned(j) * 64 + lane;
    pv[j] = valid[j] && ok ? unsigned(((kv[j] - kmin) << BQ_IDX_BITS) | p) : 0u;
    if (!ok) valid[j] = false;
  }
  // the job's group count up front when its keys span at most
  // BQ_EARLY_BITS bits (a bitmap of the keys present, in the second row
  // buffer, which the sort passes only use after it):
  // publishing it now lets later jobs' look-backs pass this job at once
  const bool early = ok && !fast && kbits <= A.early_bits;
  if (early) {
    unsigned* bits = pk(1);
    for (unsigned w = tid; w < (1u << BQ_EARLY_BITS) / 32; w += BJ_NT) bits[w] = 0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IT; j++)
      if (valid[j]) {
        const unsigned r = pv[j] >> BQ_IDX_BITS;
        atomicOr(&bits[r >> 5], 1u << (r & 31));
      }
    __syncthreads();
    unsigned c = 0;
    for (unsigned w = tid; w < (1u << BQ_EARLY_BITS) / 32; w += BJ_NT) c += unsigned(__popc(bits[w]));
    unsigned ge = 0;
    (void)block_excl_scan<BJ_NT>(c, scan, &ge);
    if (tid == 0)
      bq_lb_store(A.job_state + it.idx, (it.idx == 0 ? CP_FLAG_P : CP_FLAG_A) | (unsigned long long)ge);
  }
  // Counting sort: per-key counts by LDS atomics; one block scan gives every
  // key's bucket start and, over the keys present, the group numbers (the
  // group count is published at once); rows go to their bucket in the
  // atomics' order, and then each row to its place within the bucket: the
  // number of the bucket's packed rows smaller than its own.  Packed rows
  // are unique, so that order is the stable one, whatever order the atomics
  // took.  A bucket of more than rank_max rows (where counting costs the
  // square of its size) leaves the ordering to the LSD passes.
  unsigned cs[4] = {0u, 0u, 0u, 0u}, sv[4] = {0u, 0u, 0u, 0u}, g0 = 0, G = 0;
  bool ranked = false;
  if (fast) {
    unsigned slot[IT];
#pragma unroll
    for (int j = 0; j < IT; j++) slot[j] = valid[j] ? atomicAdd(&wcnt[pv[j] >> BQ_IDX_BITS], 1u) : 0u;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; i++) cs[i] = wcnt[4 * tid + i];
    unsigned sum = 0, nz = 0, mx = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      sum += cs[i];
      nz += cs[i] ? 1u : 0u;
      mx = cs[i] > mx ? cs[i] : mx;
    }
    // rows (at most 2^13) in the low half, keys present in the high half
    unsigned all = 0;
    const unsigned ex = block_excl_scan<BJ_NT>(sum | (nz << 16), scan, &all);
    unsigned st = ex & 0xFFFFu;
    g0 = ex >> 16;
    G = all >> 16;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      sv[i] = st;
      st += cs[i];
    }
    // (start, count) of every key
#pragma unroll
    for (int i = 0; i < 4; i++) wcnt[4 * tid + i] = sv[i] | (cs[i] << 16);
    if (mx > A.rank_max) misc[3] = 1;
    if (tid == 0)
      bq_lb_store(A.job_state + it.idx, (it.idx == 0 ? CP_FLAG_P : CP_FLAG_A) | (unsigned long long)G);
    __syncthreads();
    if (misc[3] == 0) {
      unsigned* un = pk(1);
      unsigned ent[IT];
#pragma unroll
      for (int j = 0; j < IT; j++) {
        ent[j] = wcnt[pv[j] >> BQ_IDX_BITS];  // pv is 0 for a row past the end
        if (valid[j]) un[(ent[j] & 0xFFFFu) + slot[j]] = pv[j];
      }
      if (tid < unsigned(BQ_RANK_PAD)) un[m + tid] = ~0u;  // above every packed row
      __syncthreads();
      // four bucket rows of every item per round, as many rounds as the
      // wave's largest bucket needs: a read past a bucket's end finds a
      // larger key or a sentinel, so it counts nothing
      unsigned r[IT], cmax = 0;
#pragma unroll
      for (int j = 0; j < IT; j++) {
        if (!valid[j]) ent[j] &= 0xFFFFu;  // count 0
        r[j] = 0;
        cmax = (ent[j] >> 16) > cmax ? ent[j] >> 16 : cmax;
      }
      for (unsigned i = 0; __ballot(i < cmax); i += 4) {
#pragma unroll
        for (int j = 0; j < IT; j++) {
          const unsigned* b = un + (ent[j] & 0xFFFFu) + i;
#pragma unroll
          for (int u = 0; u < 4; u++) r[j] += b[u] < pv[j] ? 1u : 0u;
        }
      }
      unsigned* out = pk(0);
#pragma unroll
      for (int j = 0; j < IT; j++)
        if (valid[j]) out[(ent[j] & 0xFFFFu) + r[j]] = pv[j];
      ranked = true;
    }
  }
  int cur = 0;
  if (!ranked) {
    // stable LSD passes in LDS
    const int npass = ok ? (kbits + A.digit_bits - 1) / A.digit_bits : 0;
    const int dbits = npass ? (kbits + npass - 1) / npass : 0;
    for (int q = 0; q < npass; q++) {
      const int sh = BQ_IDX_BITS + q * dbits;
      const unsigned dm = (1u << dbits) - 1;
      unsigned dig[IT], rank[IT];
#pragma unroll
      for (int j = 0; j < IT; j++) dig[j] = (pv[j] >> sh) & dm;
      const unsigned total = bq_rank<BJ_NT, IT>(dig, valid, dbits, wcnt, rank);
      const unsigned ex = block_excl_scan<BJ_NT>(tid < 256 ? total : 0u, scan, (unsigned*)nullptr);
      if (tid < 256) tstart[tid] = ex;
      __syncthreads();
      unsigned* out = pk(cur ^ 1);
#pragma unroll
      for (int j = 0; j < IT; j++)
        if (valid[j]) out[tstart[dig[j]] + wcnt[wave * 256 + dig[j]] + rank[j]] = pv[j];
      __syncthreads();
      cur ^= 1;
#pragma unroll
      for (int j = 0; j < IT; j++) {
        const unsigned p = wave_base + unsigned(j) * 64 + lane;
        if (valid[j]) pv[j] = out[p];
      }
    }
    if (npass == 0) {
      unsigned* srt0 = pk(0);
#pragma unroll
      for (int j = 0; j < IT; j++) {
        const unsigned p = wave_base + unsigned(j) * 64 + lane;
        if (valid[j]) srt0[p] = pv[j];
      }
    }
  }
  unsigned* srt = pk(cur);
  __syncthreads();
  // the sorted rows go out first (they need no group offset)
  const unsigned mm = ok ? unsigned(m) : 0u;
  {
    unsigned v[IT];
    double x[IT];
#pragma unroll
    for (int j = 0; j < IT; j++) {
      const unsigned i = unsigned(j) * BJ_NT + tid;
      v[j] = srt[i < mm ? i : 0u];
    }
#pragma unroll
    for (int j = 0; j < IT; j++) x[j] = vin[v[j] & (BQ_CAP - 1)];
#pragma unroll
    for (int j = 0; j < IT; j++) {
      const unsigned i = unsigned(j) * BJ_NT + tid;
      if (i < mm) {
        A.keys_out[a + i] = int64_t(kmin + (v[j] >> BQ_IDX_BITS));
        A.vals_out[a + i] = x[j];
      }
    }
  }
  unsigned* starts = pk(cur ^ 1);  // G entries (the end of the last group is m)
  if (fast) {
    // group starts from the counts
    if (wave == 0) {
      const unsigned long long off = bq_job_resolve(A.job_state, it.idx, G);
      if (lane == 0) misc[2] = off;
    }
    unsigned g = g0;
#pragma unroll
    for (int i = 0; i < 4; i++)
      if (cs[i]) starts[g++] = sv[i];
  } else {
    // group starts: a row whose key differs from the row before it
    unsigned long long fl[IT];
    unsigned wtot = 0;
#pragma unroll
    for (int j = 0; j < IT; j++) {
      const unsigned p = wave_base + unsigned(j) * 64 + lane;
      const bool st = valid[j] && (p == 0 || (srt[p] >> BQ_IDX_BITS) != (srt[p - 1] >> BQ_IDX_BITS));
      fl[j] = __ballot(st);
      wtot += unsigned(__popcll(fl[j]));
    }
    unsigned* wsum = tstart;  // per-wave group counts
    if (lane == 0) wsum[wave] = wtot;
    __syncthreads();
    unsigned before = 0;
#pragma unroll
    for (int w = 0; w < BJ_WAVES; w++) {
      const unsigned c = wsum[w];
      if (unsigned(w) < wave) before += c;
      G += c;
    }
    // group offset: decoupled look-back over the jobs (wave 0)
    if (wave == 0) {
      const unsigned long long off = early ? bq_job_resolve(A.job_state, it.idx, G)
                                           : compact_lookback(A.job_state, it.idx, G);
      if (lane == 0) misc[2] = off;
    }
    unsigned run = before;
    const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
#pragma unroll
    for (int j = 0; j < IT; j++) {
      const unsigned p = wave_base + unsigned(j) * 64 + lane;
      if ((fl[j] >> lane) & 1ull) starts[run + unsigned(__popcll(fl[j] & lt))] = p;
      run += unsigned(__popcll(fl[j]));
    }
  }
  __syncthreads();
  const unsigned long long goff = misc[2];
  if (tid == 0) {
    if (small || it.aux == 0) A.gseg_off[it.seg] = goff;
    if (it.idx == A.njobs - 1) {
      A.gseg_off[A.nseg] = goff + G;
      A.scal[1] = goff + G;
    }
  }
  for (unsigned g = tid; g < G; g += BJ_NT) {
    const unsigned s = starts[g], e = g + 1 < G ? starts[g + 1] : mm;
    double total = 0;
    // four rows per round (their loads together), added in row order
    for (unsigned i = s; i < e; i += 4) {
      unsigned v[4];
      double x[4];
#pragma unroll
      for (int u = 0; u < 4; u++) v[u] = srt[i + u < e ? i + u : e - 1];
#pragma unroll
      for (int u = 0; u < 4; u++) x[u] = vin[v[u] & (BQ_CAP - 1)];
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (i + u < e) total += x[u];
    }
    const unsigned long long o = goff + g;
    if (o < A.capacity) {
      A.gkeys[o] = int64_t(kmin + (srt[s] >> BQ_IDX_BITS));
      A.gsum[o] = 0.0 + total;
      A.gcount[o] = e - s;
    }
  }
}

// A.items holds [chunks | tiles | jobs]; the kernels get their section's
// first item
__global__ __launch_bounds__(BQ_NT) void bq_hist_kernel(BqArgs A, unsigned base) {
  extern __shared__ unsigned long long bq_lds_raw[];
  bq_hist(A, A.items[base + blockIdx.x], reinterpret_cast<unsigned char*>(bq_lds_raw));
}

__global__ __launch_bounds__(BQ_NT) void bq_plan_kernel(BqArgs A) {
  extern __shared__ unsigned long long bq_lds_raw[];
  bq_plan(A, blockIdx.x, reinterpret_cast<unsigned char*>(bq_lds_raw));
}

// tiles in arrival order (a ticket), so the look-back never waits on a tile
// that has not started
__global__ __launch_bounds__(BQ_NT) void bq_tile_kernel(BqArgs A, unsigned base) {
  extern __shared__ unsigned long long bq_lds_raw[];
  __shared__ unsigned item_s;
  if (threadIdx.x == 0) item_s = atomicAdd(A.ticket, 1u);
  __syncthreads();
  bq_tile(A, A.items[base + item_s], reinterpret_cast<unsigned char*>(bq_lds_raw));
}

// The persistent launch: each workgroup takes jobs by ticket until the
// queue ends (the next ticket is fetched while the current job runs).
__global__ __launch_bounds__(BJ_NT) void bq_jobs_kernel(BqArgs A, unsigned base, unsigned njobs) {
  extern __shared__ unsigned long long bq_lds_raw[];
  unsigned char* lds = reinterpret_cast<unsigned char*>(bq_lds_raw);
  __shared__ unsigned item_s;
  if (threadIdx.x == 0) item_s = atomicAdd(A.ticket + 1, 1u);
  for (;;) {
    __syncthreads();
    const unsigned i = item_s;
    __syncthreads();
    if (i >= njobs) break;
    unsigned next = 0;
    if (threadIdx.x == 0) next = atomicAdd(A.ticket + 1, 1u);
    bq_job(A, A.items[base + i], lds);
    if (threadIdx.x == 0) item_s = next;
  }
}

__global__ void bq_zero_kernel(unsigned long long* __restrict__ p, size_t n) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

struct Plan {
  // shape and pointers the graph was captured for
  const void *keys, *vals, *seg_off;
  size_t n, nseg;
  int key_bits;
  int cbytes;  // width of the composite key: 4 when segment + key bits fit 32
  void *keys_out, *vals_out, *gkeys, *gsum, *gcount, *gseg_off;
  size_t capacity;
  // owned
  DBuf ws, comp, comp_sorted, starts, counts, scalars;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  ~Plan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
  bool same(const Plan& o) const {
    return keys == o.keys && vals == o.vals && seg_off == o.seg_off && n == o.n &&
           nseg == o.nseg && key_bits == o.key_bits && cbytes == o.cbytes && keys_out == o.keys_out &&
           vals_out == o.vals_out && gkeys == o.gkeys && gsum == o.gsum && gcount == o.gcount &&
           gseg_off == o.gseg_off && capacity == o.capacity;
  }
};

std::unique_ptr<Plan> g_plan;  // the composite-key path

hipStream_t g_capture_stream = nullptr;

template <class CT>
void enqueue_t(Plan& p, int nbytes, hipStream_t st) {
  const size_t n = p.n;
  unsigned long long* scal = p.scalars.as<unsigned long long>();  // [0]=err [1]=G
  clear_kernel<<<1, 64, 0, st>>>(scal, 4);  // a kernel, not a memset node
  composite_kernel<CT><<<unsigned(ceil_div(n, CB_TILE)), CB_NT, 0, st>>>(
      static_cast<const int64_t*>(p.keys), n, static_cast<const size_t*>(p.seg_off), p.nseg,
      p.key_bits, p.comp.as<CT>(), reinterpret_cast<int*>(scal));
  FHIP_CHECK(hipGetLastError());
  {
    ProfScope prof("batch_sort", st);
    sort_pairs_async(int(sizeof(CT)), nbytes, p.comp.get(), p.comp_sorted.get(), 8, p.vals,
                     p.vals_out, n, p.ws.get(), st);
  }
  // run starts (device-side compaction; the count stays on the device)
  const size_t nb = ceil_div(n, CP_TILE);
  unsigned long long* counts = p.counts.as<unsigned long long>();
  RunStart<CT> pred{p.comp_sorted.as<CT>()};
  compact_count_kernel<<<unsigned(nb), CP_NT, 0, st>>>(pred, n, counts);
  scan_counts_kernel<<<1, 1024, 0, st>>>(counts, nb, scal + 1);
  compact_write_kernel<<<unsigned(nb), CP_NT, 0, st>>>(pred, IdentityMap{}, n, counts,
                                                       p.starts.as<size_t>());
  FHIP_CHECK(hipGetLastError());
  {
    ProfScope prof("batch_fold", st);
    fold_kernel<CT><<<grid_for_n(n), 256, 0, st>>>(
        p.comp_sorted.as<CT>(), static_cast<const double*>(p.vals_out), p.starts.as<size_t>(),
        scal + 1, n, p.key_bits, p.capacity, static_cast<int64_t*>(p.gkeys),
        static_cast<double*>(p.gsum), static_cast<uint64_t*>(p.gcount));
    FHIP_CHECK(hipGetLastError());
  }
  strip_kernel<CT><<<grid_for_n(n), 256, 0, st>>>(p.comp_sorted.as<CT>(), n, p.key_bits,
                                                  static_cast<int64_t*>(p.keys_out));
  FHIP_CHECK(hipGetLastError());
  seg_groups_kernel<CT><<<grid_for_n(n + 1), 256, 0, st>>>(
      p.comp_sorted.as<CT>(), p.starts.as<size_t>(), scal + 1, p.nseg, p.key_bits,
      static_cast<size_t*>(p.gseg_off));
  FHIP_CHECK(hipGetLastError());
}

void enqueue(Plan& p, hipStream_t st) {
  const int total_bits = p.key_bits + (p.nseg > 1 ? 64 - __builtin_clzll(p.nseg - 1) : 0);
  const int nbytes = std::max(1, (total_bits + 7) / 8);
  if (p.cbytes == 4) enqueue_t<uint32_t>(p, nbytes, st);
  else enqueue_t<uint64_t>(p, nbytes, st);
}

}  // namespace

void composite_sort_groupby(const void* keys, const void* vals, size_t n, const size_t* seg_off,
                        size_t nseg, int key_bits, void* keys_out, void* vals_out, void* gkeys,
                        void* gsum, void* gcount, size_t* gseg_off, size_t capacity,
                        size_t* ngroups, int use_graph, hipStream_t st) {
  FHIP_REQUIRE(key_bits >= 1 && key_bits <= 48, FHIP_ERR_ARG, "batch: key_bits must be 1..48");
  FHIP_REQUIRE(nseg >= 1 && nseg < (size_t(1) << 20), FHIP_ERR_ARG, "batch: 1 to 2^20 segments");
  *ngroups = 0;
  if (n == 0) {
    std::vector<size_t> z(nseg + 1, 0);
    FHIP_CHECK(hipMemcpyAsync(gseg_off, z.data(), (nseg + 1) * sizeof(size_t),
                              hipMemcpyHostToDevice, st));
    FHIP_CHECK(hipStreamSynchronize(st));
    return;
  }
  auto p = std::make_unique<Plan>();
  p->keys = keys; p->vals = vals; p->seg_off = seg_off; p->n = n; p->nseg = nseg;
  p->key_bits = key_bits; p->keys_out = keys_out; p->vals_out = vals_out; p->gkeys = gkeys;
  p->gsum = gsum; p->gcount = gcount; p->gseg_off = gseg_off; p->capacity = capacity;
  const int seg_bits = nseg > 1 ? 64 - __builtin_clzll(nseg - 1) : 0;
  FHIP_REQUIRE(key_bits + seg_bits <= 64, FHIP_ERR_ARG, "batch: key_bits + segment bits > 64");
  p->cbytes = key_bits + seg_bits <= 32 ? 4 : 8;
  Plan* plan = p.get();
  if (use_graph && g_plan && g_plan->same(*p)) {
    plan = g_plan.get();
  } else {
    p->ws.reset(sort_async_workspace_bytes(n, p->cbytes, 8));
    p->comp.reset(n * size_t(p->cbytes));
    p->comp_sorted.reset(n * size_t(p->cbytes));
    p->starts.reset(n * sizeof(size_t));
    p->counts.reset((ceil_div(n, CP_TILE) + 1) * 8);
    p->scalars.reset(4 * 8);
    if (use_graph) {
      if (!g_capture_stream) FHIP_CHECK(hipStreamCreateWithFlags(&g_capture_stream, hipStreamNonBlocking));
      FHIP_CHECK(hipStreamBeginCapture(g_capture_stream, hipStreamCaptureModeThreadLocal));
      try {
        enqueue(*p, g_capture_stream);
      } catch (...) {
        hipGraph_t dropped = nullptr;
        (void)hipStreamEndCapture(g_capture_stream, &dropped);
        if (dropped) (void)hipGraphDestroy(dropped);
        throw;
      }
      FHIP_CHECK(hipStreamEndCapture(g_capture_stream, &p->graph));
      FHIP_CHECK(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
      g_plan = std::move(p);
      plan = g_plan.get();
    }
  }
  if (use_graph) {
    FHIP_CHECK(hipGraphLaunch(plan->exec, st));
  } else {
    enqueue(*plan, st);
  }
  unsigned long long scal[2] = {0, 0};
  FHIP_CHECK(hipMemcpyAsync(scal, plan->scalars.get(), sizeof(scal), hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  FHIP_REQUIRE(unsigned(scal[0]) == 0, FHIP_ERR_ARG,
               "batch: a key is outside [0, 2^key_bits)");
  *ngroups = size_t(scal[1]);
  FHIP_REQUIRE(*ngroups <= capacity, FHIP_ERR_TOO_LARGE, "batch: group capacity too small");
}


namespace {

// Queue and buffers of the work-queue path for one shape (and one content
// of seg_off, which the queue is built from); the HIP graph replays it.
struct QPlan {
  const void *keys = nullptr, *vals = nullptr, *seg_off = nullptr;
  size_t n = 0, nseg = 0;
  int key_bits = 0;
  void *keys_out = nullptr, *vals_out = nullptr, *gkeys = nullptr, *gsum = nullptr,
       *gcount = nullptr, *gseg_off = nullptr;
  size_t capacity = 0;
  std::vector<size_t> off;  // seg_off contents
  long long tunables[5] = {0, 0, 0, 0, 0};
  std::vector<long long> class_rows;  // batch.bucket_rows.c<k> for every k (-1: unset)
  unsigned nitems = 0, njobs = 0, nlarge = 0, nchunks = 0, ntiles = 0;
  size_t lb_words = 0;
  DBuf items, lsegs, fine_hist, bmap, bstart, lookback, job_state, tk, tv, scal;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  ~QPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
  bool same(const QPlan& o) const {
    return keys == o.keys && vals == o.vals && seg_off == o.seg_off && n == o.n &&
           nseg == o.nseg && key_bits == o.key_bits && keys_out == o.keys_out &&
           vals_out == o.vals_out && gkeys == o.gkeys && gsum == o.gsum && gcount == o.gcount &&
           gseg_off == o.gseg_off && capacity == o.capacity && off == o.off &&
           std::equal(tunables, tunables + 5, o.tunables) && class_rows == o.class_rows;
  }
};

std::unique_ptr<QPlan> g_qplan;

// Rows per bucket job of a large segment of m rows: the tunable of its size
// class (log2 of m, rounded up) if set, else "batch.bucket_rows".
unsigned long long bucket_rows_for(size_t m, long long dflt) {
  int c = 0;
  while ((size_t(1) << c) < m) c++;
  const std::string name = "batch.bucket_rows.c" + std::to_string(c);
  const long long v = tune_get(name.c_str(), dflt);
  return (unsigned long long)std::max(256LL, std::min(v, (long long)BQ_CAP));
}

void build_queue(QPlan& q, hipStream_t st) {
  const long long bucket_rows = tune_get("batch.bucket_rows", 7168);
  const long long chunk = std::max(1024LL, tune_get("batch.hist_chunk", 1 << 16));
  const int fshift = std::max(0, q.key_bits - BQ_FINE_BITS);
  std::vector<BqSeg> ls;
  std::vector<BqItem> hist, tiles, jobs;
  unsigned chunk0 = 0, tile0 = 0, job = 0;
  for (size_t s = 0; s < q.nseg; s++) {
    const unsigned long long a = q.off[s], b = q.off[s + 1], m = b - a;
    if (m <= (unsigned long long)BQ_CAP) {
      jobs.push_back(BqItem{BQ_JOB, unsigned(s), job++, BQ_SMALL, a, b});
      continue;
    }
    const unsigned L = unsigned(ls.size());
    BqSeg g{};
    g.start = a;
    g.len = m;
    g.chunk0 = chunk0;
    g.nchunks = unsigned(ceil_div(m, size_t(chunk)));
    g.tile0 = tile0;
    g.ntiles = unsigned(ceil_div(m, size_t(BQ_TILE_ROWS)));
    g.nb = unsigned(std::min<unsigned long long>(256, ceil_div(m, bucket_rows_for(m, bucket_rows))));
    g.fshift = fshift;
    for (unsigned c = 0; c < g.nchunks; c++)
      hist.push_back(BqItem{BQ_HIST, L, chunk0 + c, 0, a + c * chunk,
                            std::min<unsigned long long>(b, a + (c + 1) * chunk)});
    for (unsigned t = 0; t < g.ntiles; t++)
      tiles.push_back(BqItem{BQ_TILE, L, t, 0, a + size_t(t) * BQ_TILE_ROWS,
                             std::min<unsigned long long>(b, a + size_t(t + 1) * BQ_TILE_ROWS)});
    for (unsigned k = 0; k < g.nb; k++) jobs.push_back(BqItem{BQ_JOB, unsigned(s), job++, k, L, 0});
    chunk0 += g.nchunks;
    tile0 += g.ntiles;
    ls.push_back(g);
  }
  std::vector<BqItem> all;
  all.reserve(hist.size() + tiles.size() + jobs.size());
  for (auto* v : {&hist, &tiles, &jobs}) all.insert(all.end(), v->begin(), v->end());
  q.nitems = unsigned(all.size());
  q.nchunks = chunk0;
  q.ntiles = tile0;
  q.njobs = job;
  q.nlarge = unsigned(ls.size());
  q.lb_words = size_t(tile0) * 256;
  q.items.reset(std::max<size_t>(1, all.size()) * sizeof(BqItem));
  q.lsegs.reset(std::max<size_t>(1, ls.size()) * sizeof(BqSeg));
  q.fine_hist.reset(std::max<size_t>(1, chunk0) * BQ_FINE * 4);
  q.bmap.reset(std::max<size_t>(1, ls.size()) * BQ_FINE);
  q.bstart.reset(std::max<size_t>(1, ls.size()) * 257 * 4);
  q.lookback.reset(std::max<size_t>(1, q.lb_words) * 8);
  q.job_state.reset(std::max<size_t>(1, q.njobs) * 8);
  q.tk.reset(std::max<size_t>(1, q.n) * 8);
  q.tv.reset(std::max<size_t>(1, q.n) * 8);
  q.scal.reset(64);
  // stream-ordered after the earlier users of these pool blocks
  FHIP_CHECK(hipMemcpyAsync(q.items.get(), all.data(), all.size() * sizeof(BqItem),
                            hipMemcpyHostToDevice, st));
  if (!ls.empty())
    FHIP_CHECK(hipMemcpyAsync(q.lsegs.get(), ls.data(), ls.size() * sizeof(BqSeg),
                              hipMemcpyHostToDevice, st));
  FHIP_CHECK(hipStreamSynchronize(st));  // the host vectors go out of scope
}

void enqueue_queue(QPlan& q, hipStream_t st) {
  unsigned long long* scal = q.scal.as<unsigned long long>();  // [0..2] scalars, [4] ticket
  bq_zero_kernel<<<1, 64, 0, st>>>(scal, 8);
  bq_zero_kernel<<<grid_for_n(q.lb_words), 256, 0, st>>>(q.lookback.as<unsigned long long>(), q.lb_words);
  bq_zero_kernel<<<grid_for_n(size_t(q.njobs)), 256, 0, st>>>(q.job_state.as<unsigned long long>(), q.njobs);
  FHIP_CHECK(hipGetLastError());
  BqArgs A{};
  A.keys = static_cast<const int64_t*>(q.keys);
  A.vals = static_cast<const double*>(q.vals);
  A.items = q.items.as<BqItem>();
  A.njobs = q.njobs;
  A.lsegs = q.lsegs.as<BqSeg>();
  A.fine_hist = q.fine_hist.as<unsigned>();
  A.bmap = q.bmap.as<unsigned char>();
  A.bstart = q.bstart.as<unsigned>();
  A.lookback = q.lookback.as<unsigned long long>();
  A.job_state = q.job_state.as<unsigned long long>();
  A.tk = q.tk.as<int64_t>();
  A.tv = q.tv.as<double>();
  A.keys_out = static_cast<int64_t*>(q.keys_out);
  A.vals_out = static_cast<double*>(q.vals_out);
  A.gkeys = static_cast<int64_t*>(q.gkeys);
  A.gsum = static_cast<double*>(q.gsum);
  A.gcount = static_cast<uint64_t*>(q.gcount);
  A.gseg_off = static_cast<size_t*>(q.gseg_off);
  A.capacity = q.capacity;
  A.nseg = q.nseg;
  A.key_lim = q.key_bits >= 64 ? ~0ull : (1ull << q.key_bits);
  A.digit_bits = int(q.tunables[0]);
  A.early_bits = int(q.tunables[3]);
  A.rank_max = unsigned(q.tunables[4]);
  A.ticket = reinterpret_cast<unsigned*>(scal + 4);
  A.scal = scal;
  {
    ProfScope prof("batch_tiles", st);
    if (q.nchunks) bq_hist_kernel<<<q.nchunks, BQ_NT, BH_LDS, st>>>(A, 0);
    if (q.nlarge) bq_plan_kernel<<<q.nlarge, BQ_NT, BP_LDS, st>>>(A);
    if (q.ntiles) bq_tile_kernel<<<q.ntiles, BQ_NT, BT_LDS, st>>>(A, q.nchunks);
    FHIP_CHECK(hipGetLastError());
  }
  ProfScope prof("batch_jobs", st);
  const unsigned grid = std::max(1u, std::min(q.njobs, unsigned(fill_blocks(BJ_NT, 1))));
  bq_jobs_kernel<<<grid, BJ_NT, BJ_LDS, st>>>(A, q.nchunks + q.ntiles, q.njobs);
  FHIP_CHECK(hipGetLastError());
}

}  // namespace

// The work-queue path; returns false when a bucket overflowed its job (the
// caller then runs the composite path).  Errors throw.
bool queue_sort_groupby(const void* keys, const void* vals, size_t n, const size_t* seg_off,
                        size_t nseg, int key_bits, void* keys_out, void* vals_out, void* gkeys,
                        void* gsum, void* gcount, size_t* gseg_off, size_t capacity,
                        size_t* ngroups, int use_graph, hipStream_t st) {
  static bool attr = false;
  if (!attr) {
    FHIP_CHECK(hipFuncSetAttribute((const void*)bq_tile_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, BT_LDS));
    FHIP_CHECK(hipFuncSetAttribute((const void*)bq_jobs_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, BJ_LDS));
    attr = true;
  }
  auto q = std::make_unique<QPlan>();
  q->keys = keys; q->vals = vals; q->seg_off = seg_off; q->n = n; q->nseg = nseg;
  q->key_bits = key_bits; q->keys_out = keys_out; q->vals_out = vals_out; q->gkeys = gkeys;
  q->gsum = gsum; q->gcount = gcount; q->gseg_off = gseg_off; q->capacity = capacity;
  q->off.resize(nseg + 1);
  FHIP_CHECK(hipMemcpyAsync(q->off.data(), seg_off, (nseg + 1) * sizeof(size_t),
                            hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  FHIP_REQUIRE(q->off[0] == 0 && q->off[nseg] == n, FHIP_ERR_ARG,
               "batch: seg_off must run from 0 to n");
  for (size_t s = 0; s < nseg; s++)
    FHIP_REQUIRE(q->off[s] <= q->off[s + 1], FHIP_ERR_ARG, "batch: seg_off must not decrease");
  q->tunables[0] = std::max(2LL, std::min(8LL, tune_get("batch.digit_bits", 8)));
  q->tunables[1] = tune_get("batch.bucket_rows", 7168);
  q->tunables[2] = tune_get("batch.hist_chunk", 1 << 16);
  q->tunables[3] = std::max(-1LL, std::min((long long)BQ_EARLY_BITS, tune_get("batch.early_bits", BQ_EARLY_BITS)));
  q->tunables[4] = std::max(0LL, std::min((long long)BQ_RANK_MAX, tune_get("batch.rank_max", BQ_RANK_MAX)));
  for (int c = 0; c < 64; c++)
    q->class_rows.push_back(tune_get(("batch.bucket_rows.c" + std::to_string(c)).c_str(), -1));
  if (!(g_qplan && g_qplan->same(*q))) {
    build_queue(*q, st);
    g_qplan = std::move(q);
  }
  QPlan* plan = g_qplan.get();
  if (use_graph && !plan->exec) {  // captured once per plan, replayed after
    if (!g_capture_stream) FHIP_CHECK(hipStreamCreateWithFlags(&g_capture_stream, hipStreamNonBlocking));
    FHIP_CHECK(hipStreamBeginCapture(g_capture_stream, hipStreamCaptureModeThreadLocal));
    try {
      enqueue_queue(*plan, g_capture_stream);
    } catch (...) {
      hipGraph_t dropped = nullptr;
      (void)hipStreamEndCapture(g_capture_stream, &dropped);
      if (dropped) (void)hipGraphDestroy(dropped);
      throw;
    }
    FHIP_CHECK(hipStreamEndCapture(g_capture_stream, &plan->graph));
    FHIP_CHECK(hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0));
  }
  if (use_graph) FHIP_CHECK(hipGraphLaunch(plan->exec, st));
  else enqueue_queue(*plan, st);
  unsigned long long sc[3] = {0, 0, 0};
  FHIP_CHECK(hipMemcpyAsync(sc, plan->scal.get(), sizeof(sc), hipMemcpyDeviceToHost, st));
  FHIP_CHECK(hipStreamSynchronize(st));
  FHIP_REQUIRE(sc[0] == 0, FHIP_ERR_ARG, "batch: a key is outside [0, 2^key_bits)");
  if (sc[2]) return false;
  *ngroups = size_t(sc[1]);
  FHIP_REQUIRE(*ngroups <= capacity, FHIP_ERR_TOO_LARGE, "batch: group capacity too small");
  return true;
}

void batch_sort_groupby(const void* keys, const void* vals, size_t n, const size_t* seg_off,
                        size_t nseg, int key_bits, void* keys_out, void* vals_out, void* gkeys,
                        void* gsum, void* gcount, size_t* gseg_off, size_t capacity,
                        size_t* ngroups, int use_graph, hipStream_t st) {
  FHIP_REQUIRE(key_bits >= 1 && key_bits <= 48, FHIP_ERR_ARG, "batch: key_bits must be 1..48");
  FHIP_REQUIRE(nseg >= 1 && nseg < (size_t(1) << 20), FHIP_ERR_ARG, "batch: 1 to 2^20 segments");
  *ngroups = 0;
  // "batch.path": 0 (default) the work queue, 1 the composite-key sort
  if (n > 0 && tune_get("batch.path", 0) == 0 &&
      queue_sort_groupby(keys, vals, n, seg_off, nseg, key_bits, keys_out, vals_out, gkeys, gsum,
                         gcount, gseg_off, capacity, ngroups, use_graph, st))
    return;
  composite_sort_groupby(keys, vals, n, seg_off, nseg, key_bits, keys_out, vals_out, gkeys, gsum,
                         gcount, gseg_off, capacity, ngroups, use_graph, st);
}

void batch_release() {
  g_plan.reset();
  g_qplan.reset();
}

}  // namespace fhip

extern "C" int fhip_dev_batch_sort_groupby(const int64_t* keys, const double* vals, size_t n,
                                           const size_t* seg_off, size_t nseg, int key_bits,
                                           int64_t* keys_out, double* vals_out, int64_t* gkeys,
                                           double* gsum, uint64_t* gcount, size_t* gseg_off,
                                           size_t capacity, size_t* ngroups, int use_graph,
                                           void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(ngroups && seg_off && gseg_off, FHIP_ERR_ARG, "batch: null pointer");
    FHIP_REQUIRE(n == 0 || (keys && vals && keys_out && vals_out && gkeys && gsum && gcount),
                 FHIP_ERR_ARG, "batch: null pointer");
    FHIP_REQUIRE(keys_out != keys && vals_out != vals, FHIP_ERR_ARG,
                 "batch: outputs must not alias the inputs");
    fhip::batch_sort_groupby(keys, vals, n, seg_off, nseg, key_bits, keys_out, vals_out, gkeys,
                             gsum, gcount, gseg_off, capacity, ngroups, use_graph,
                             fhip::as_stream(stream));
  });
}
