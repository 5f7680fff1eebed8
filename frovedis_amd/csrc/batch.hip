// Batched segments: thousands of independent (sort + group sum/count) jobs
// on one device in one HIP graph launch.  No counterpart in the reference,
// which runs each job as radix_sort(key, val) (core/radix_sort.hpp:197-219)
// followed by the per-group sequential fold of dfcolumn_grouped_agg.cc:
// 153-164; the outputs here are those of that per-segment loop.
//
// Segment s holds rows [seg_off[s], seg_off[s+1]) with keys in
// [0, 2^key_bits).  Every row gets the composite key (s << key_bits) | key,
// one stable LSD sort orders all segments at once (segment-major, then key,
// ties in input order), and one thread per group folds its run of values in
// order -- the same additions in the same order as the reference, so the
// sums are bit-exact.  The whole sequence -- composite keys, histogram,
// passes, run detection, fold -- is stream-ordered with no host
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

std::unique_ptr<Plan> g_plan;
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

void batch_sort_groupby(const void* keys, const void* vals, size_t n, const size_t* seg_off,
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

void batch_release() { g_plan.reset(); }

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
