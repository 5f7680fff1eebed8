/*
 * libfrovedis_hip: C ABI of the MI355X (gfx950) dataframe execution core.
 *
 * The entry points replace, for one rank's local partition, the reference's
 * header templates and per-rank lambdas (paths relative to the frovedis
 * source tree, src/frovedis/).  The C++ shims under include/frovedis_hip/
 * keep the reference's own signatures and forward here.
 *
 * Conventions
 *  - Every entry point returns an int status (FHIP_OK == 0).  On failure
 *    fhip_last_error() returns the message; the C++ shims rethrow it as
 *    std::runtime_error, as the reference does.
 *  - fhip_dev_* entry points take device pointers and a hipStream_t passed as
 *    void*; NULL means the default stream.  They may synchronise that stream
 *    when a result size must reach the host.
 *  - Entry points without the dev_ prefix take host pointers (the reference's
 *    std::vector storage), stage through pinned memory and return when done.
 *  - Calls on one device must not overlap (one host thread per rank, like the
 *    reference's MPI_THREAD_SERIALIZED ranks).
 */
#ifndef FROVEDIS_HIP_H
#define FROVEDIS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define FHIP_OK 0
#define FHIP_ERR_ARG 1
#define FHIP_ERR_HIP 2
#define FHIP_ERR_NOMEM 3
#define FHIP_ERR_TYPE 4
#define FHIP_ERR_INTERNAL 5
#define FHIP_ERR_TOO_LARGE 6

/* element type codes */
#define FHIP_I32 0
#define FHIP_U32 1
#define FHIP_I64 2
#define FHIP_U64 3
#define FHIP_F32 4
#define FHIP_F64 5

/* ---- runtime ------------------------------------------------------------ */
/* Selects the device and warms the runtime.  Optional: every entry point
 * works on the current device without it.  No reference counterpart
 * (frovedis_init.hpp:9-19 use_frovedis is the closest). */
int fhip_init(int device);
/* Frees cached device workspace. */
int fhip_finalize(void);
/* Message of the last failed call on this thread. */
const char* fhip_last_error(void);
/* Frees device memory that an entry point allocated for its output. */
int fhip_dev_free(void* p);

/* Kernel timing (the role of the reference's time_spent laps,
 * core/utility.hpp:44-75): when enabled, the library records HIP events on
 * the launch stream around its main kernels, named "radix_pass",
 * "radix_hist", "hashagg", "ht_build", "ht_probe", "join_probe",
 * "batch_sort", "learned_scatter"...  fhip_prof_enable(on) also clears
 * what was recorded; fhip_prof_query sums the durations of one name. */
int fhip_prof_enable(int on);
int fhip_prof_query(const char* name, double* total_ms, size_t* launches);

/* ---- radix_sort ---------------------------------------------------------
 * Stable ascending (or descending) sort of keys with an optional payload,
 * in place.  Replaces radix_sort(K*, V*, size_t, bool positive_only)
 * (core/radix_sort.hpp:197-219), its key-only form (:221-242), the float and
 * double overloads (:244-340), radix_sort_desc (:342-487) and the vector
 * forms (:489-518).  vbytes is 0 (no payload), 4 or 8. */
int fhip_radix_sort(int ktype, void* keys, int vbytes, void* vals, size_t n,
                    int positive_only, int descending);
int fhip_dev_radix_sort(int ktype, void* keys, int vbytes, void* vals,
                        size_t n, int positive_only, int descending,
                        void* stream);

/* ---- prefix_sum ---------------------------------------------------------
 * Inclusive scan, prefix_sum(const T*, T*, size_t) (core/prefix_sum.hpp:
 * 67-88).  Integer types are exact; float types are summed in a fixed tree
 * order (not the reference's sequential order). */
int fhip_prefix_sum(int type, const void* in, void* out, size_t n);
int fhip_dev_prefix_sum(int type, const void* in, void* out, size_t n,
                        void* stream);

/* ---- set_separate / set_is_unique ----------------------------------------
 * Run boundaries of a sorted array: [0, i1, ..., n], or {0} when n == 0
 * (core/set_operations.hpp:146-163).  `out` holds n + 1 entries; *out_n
 * (host) receives the number written. */
int fhip_set_separate(int type, const void* sorted, size_t n, size_t* out,
                      size_t* out_n);
int fhip_dev_set_separate(int type, const void* sorted, size_t n,
                          size_t* out, size_t* out_n, void* stream);
/* set_is_unique (core/set_operations.hpp:184-197) of a sorted array. */
int fhip_dev_set_is_unique(int type, const void* sorted, size_t n,
                           int* is_unique, void* stream);

/* ---- lower_bound / upper_bound -------------------------------------------
 * Batched binary search (core/lower_bound.hpp:13-101,
 * core/upper_bound.hpp). */
int fhip_dev_lower_bound(int type, const void* sorted, size_t m,
                         const void* queries, size_t nq, size_t* out,
                         int upper, void* stream);

/* ---- group_by + grouped aggregation (sort path) ---------------------------
 * typed_dfcolumn<T>::group_by on one rank (dataframe/dfcolumn_groupby.cc:
 * 16-120): gathers col[idx[i]], stable-sorts (key, idx) -- idx is permuted
 * in place into local_grouped_idx -- and writes local_idx_split (n + 1
 * entries of capacity; *nsplit receives the count) and the group keys
 * (ascending, nsplit - 1 of them). */
int fhip_dev_group_by(int ktype, const void* col, size_t* idx, size_t n,
                      size_t* split, size_t* nsplit, void* keys_out,
                      void* stream);

/* aggregation codes (the dfaggregator sum/count/avg/max/min,
 * dataframe/dfaggregator.cc:6-63) */
#define FHIP_AGG_SUM 0
#define FHIP_AGG_COUNT 1
#define FHIP_AGG_AVG 2
#define FHIP_AGG_MAX 3
#define FHIP_AGG_MIN 4

/* Per-group aggregate of one column over (grouped_idx, split) from
 * fhip_dev_group_by (dataframe/dfcolumn_grouped_agg.cc: sum :105-275,
 * max :278-510, min :513-717, count :719-837, avg :839-889; one rank).
 * `nulls` lists null row positions (may be NULL when nnull == 0).  Output
 * types: sum/max/min in vtype, count as uint64, avg as double. */
int fhip_dev_grouped_agg(int op, int vtype, const void* col,
                         const size_t* grouped_idx, size_t n,
                         const size_t* split, size_t nsplit,
                         const size_t* nulls, size_t nnull, void* out,
                         void* stream);

/* ---- fused group_by + aggregation (hash path) ----------------------------
 * One pass over the rows for group_by(k).select({k}, {aggs...})
 * (dftable.cc:346-366 and :3724-3772 on one rank, no nulls): integer key
 * column, up to four distinct (column, sum|max|min) accumulators plus the
 * row count.  Groups come out ascending by key; keys_out and every
 * aggs[i].out hold `capacity` entries; *ngroups receives the group count.
 * table_hint (0 = automatic) is the expected number of groups. */
typedef struct fhip_agg {
  int op;          /* FHIP_AGG_* */
  int vtype;       /* FHIP_* type code of col (ignored for count) */
  const void* col; /* device column of n rows (NULL for count) */
  void* out;       /* device output */
} fhip_agg;
int fhip_dev_group_by_agg(int ktype, const void* keys, size_t n,
                          const fhip_agg* aggs, int naggs, void* keys_out,
                          size_t capacity, size_t* ngroups, size_t table_hint,
                          void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FROVEDIS_HIP_H */
