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
 * what was recorded; fhip_prof_query sums the durations of one name.
 * fhip_prof_filter(names) records only the comma-separated names given
 * (NULL or "" records every name), so that timing one kernel costs a
 * timed region two event records per launch of it. */
int fhip_prof_enable(int on);
int fhip_prof_filter(const char* names);
int fhip_prof_query(const char* name, double* total_ms, size_t* launches);

/* Tunable parameters, set by the autotuner (frovedis_amd/autotune.py) from
 * measurements on the device and kept in a JSON cache:
 *   "radix.shape_k", "radix.shape_kv"  tile shape of the radix sort pass
 *       (index: 0 256x16, 1 256x32, 2 512x16, 3 512x24, 4 512x12, 5 1024x8,
 *       6 1024x16, 7 768x16 threads x keys) for key-only / key-value sorts;
 *   "radix.shape_narrow"  tile shape of the sort-merge join's narrow sort
 *       (32-bit keys and rows): 2, 3 (default) or 6;
 *   "join.narrow"       0 switches off the sort-merge join's narrow path;
 *   "groupby.cold32"    0 keeps whole 64-bit cold keys in the group-by's
 *       heavy-key pass (else their low 32 bits when the range path applies);
 *   "groupby.slots"     LDS hash table slots of the fused group-by (power of 2);
 *   "groupby.fill_pct"  expected groups per partition as a percentage of the
 *       slots (the hash load factor);
 *   "groupby.chunk"     rows aggregated by one workgroup;
 *   "groupby.span"      4096-row tiles partitioned by one workgroup;
 *   "groupby.hot"       slots of the heavy-key table (0 switches the heavy-key
 *       pass off); "groupby.hot_min_pct" the sample share the heavy keys
 *       must cover for the pass to run; "groupby.hot_sample" log2 of the
 *       rows sampled to find them (12 to 20, default 17);
 *   "compact.shape"     tile of the one-pass stream compaction behind filter,
 *       set_separate and the join outputs (0 256x16, 1 1024x16 threads x
 *       rows);
 *   "multi_group_by.composite"  0 switches off the composite-key path of
 *       fhip_dev_multi_group_by (the column-by-column sorts run instead);
 *   "join.load_pct"     load factor (10 to 50 %, default 50) of the joins' hash table for
 *       integer keys;
 *   "join.dense_min", "join.dense_ratio"  the joins' direct table for integer
 *       build keys that are unique and fill their range: at least this many
 *       build rows (default 2^20) and a key range of at most dense_ratio x
 *       the rows (default 2; 0 switches the direct table off);
 *   "join.dense_slice_bits", "join.dense_chunks"  key offsets per part of the
 *       direct table's radix partitioning (2^bits, 10 to 18, default 18) and
 *       the workgroups that probe one part (default 128);
 *       "join.dense_lds_build" 0 builds a table of row positions by scatter
 *       and check instead of in LDS parts of 2^15 offsets (default 1);
 *       "join.dense_probe_u" rows per lane in flight in its probe (4, 8, 16
 *       or 32, default 32);
 *   "groupby.range", "groupby.range_max", "groupby.range_chunks"  0 switches
 *       off the fused group-by's range path (keys within a narrow range:
 *       parts by key offset, LDS arrays indexed by it); the widest key range
 *       it takes (at most 2^21, and at most the row count); workgroups per
 *       part of its aggregation (default: enough for 512 workgroups);
 *   "radix.hist_bpc", "radix.hist_mode"  workgroups per CU of the radix
 *       histogram pass (1 to 64; by default about 16384 keys per workgroup,
 *       8 to 64 workgroups per CU), and how a wave counts its digits: one
 *       atomic for the lanes holding lane 0's digit when at least 8 do, else
 *       one per lane, the match found from lane 0's whole key (4, default)
 *       or its digit per byte (2); 3 one atomic per lane, no match;
 *   "groupby.dense_ratio"  the group keys are ordered by ranks from flags
 *       and a scan, not sorted, when their range holds at most this many
 *       possible keys per group (default 32; 0 always sorts);
 *   "batch.path"        0 (default) the batched work queue, 1 the
 *       composite-key sort of fhip_dev_batch_sort_groupby;
 *   "batch.digit_bits"  bits per LSD pass of a job's LDS sort (2 to 8,
 *       default 8);
 *   "batch.bucket_rows" rows per bucket job of a segment larger than one
 *       job (8192 rows; default 7168), "batch.bucket_rows.c<k>" the same
 *       for segments of 2^(k-1) to 2^k rows;
 *   "batch.hist_chunk", "batch.early_bits"  rows per histogram workgroup of
 *       the large segments (default 65536) and the key span (bits, default
 *       15) up to which a job publishes its group count before sorting;
 *   "batch.rank_max"    a job whose keys span at most 4096 values is sorted
 *       by counting when no key has more than this many rows (default 64;
 *       0: always the LSD passes).
 * A negative value removes the setting (back to the built-in default).  No
 * counterpart in the reference. */
int fhip_tune_set(const char* name, long long value);
int fhip_tune_get(const char* name, long long dflt, long long* value);
/* The tunables as one flat JSON object {"name": integer, ...}: save writes
 * every value that is set; load sets every value in the file. */
int fhip_autotune_save(const char* path);
int fhip_autotune_load(const char* path);

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
/* No counterpart in the reference: fhip_dev_radix_sort for callers that
 * cannot allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results; it calls no hipMalloc /
 * hipFree: its temporaries live in a file-scope 64 MiB device array, one
 * per GPU, kept until the process ends.  Like fhip_dev_radix_sort it reads
 * the digit histogram back once (a wait on the stream) to skip the bytes
 * every key shares, so it is not meant for graph capture.  Inputs whose temporaries do not
 * fit (more than about 4M 8+8-byte pairs, 8M 8-byte keys) return
 * FHIP_ERR_TOO_LARGE.  Every call on a GPU shares that array: calls that
 * overlap on one GPU (two streams at once, two threads) race. */
int fhip_dev_radix_sort_noalloc(int ktype, void* keys, int vbytes, void* vals,
                                size_t n, int positive_only, int descending,
                                void* stream);

/* ---- prefix_sum ---------------------------------------------------------
 * Inclusive scan, prefix_sum(const T*, T*, size_t) (core/prefix_sum.hpp:
 * 67-88).  Integer types are exact; float types are summed in a fixed tree
 * order (not the reference's sequential order).
 * Host-pointer forms (no fhip_dev_ prefix) run on the default stream and
 * copy through pinned staging buffers of the calling thread; every entry
 * point orders its stream after the previous call's stream on the device,
 * so workspace blocks reused across streams never overlap in time. */
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
/* set_is_unique (core/set_operations.hpp:184-197) of a sorted array; the
 * host form copies the keys to the device and runs the same kernel. */
int fhip_set_is_unique(int type, const void* sorted, size_t n, int* is_unique);
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
 * col holds col_n rows; grouped_idx (n entries) indexes it.  `nulls` lists
 * null row positions of the column (may be NULL when nnull == 0).  Output
 * types: sum/max/min in vtype, count as uint64, avg as double. */
int fhip_dev_grouped_agg(int op, int vtype, const void* col, size_t col_n,
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

/* ---- unique_hashtable<K, size_t> -----------------------------------------
 * Same layout and results as dataframe/hashtable.hpp: table size = smallest
 * listed prime >= 3n, home slot = key bits % size, last row owns a slot,
 * other keys go to a sorted conflict list.  Build replaces the (k, v) and
 * (k, v, int& is_unique_ok) constructors (:1167-1322; vals NULL gives the
 * set constructor :1324-1384); lookup replaces lookup(k, miss) (:1442-1515:
 * values of missed keys are 0, misses are returned as sorted row indices);
 * check_existence (:1517-1582); all_keys (:1584-1607: owned slots in slot
 * order, then the conflict keys).  miss_out and all_keys' out hold n (resp.
 * build n) entries. */
typedef struct fhip_ht fhip_ht;
int fhip_dev_ht_build(int ktype, const void* keys, const uint64_t* vals,
                      size_t n, int check_unique, int* is_unique_ok,
                      fhip_ht** ht, void* stream);
int fhip_dev_ht_lookup(const fhip_ht* ht, const void* keys, size_t n,
                       uint64_t* vals_out, size_t* miss_out, size_t* miss_n,
                       void* stream);
int fhip_dev_ht_check_existence(const fhip_ht* ht, const void* keys, size_t n,
                                int* out, void* stream);
int fhip_dev_ht_all_keys(const fhip_ht* ht, void* out, size_t* out_n,
                         void* stream);
int fhip_ht_info(const fhip_ht* ht, size_t* table_size, size_t* conflict_size);
int fhip_ht_destroy(fhip_ht* ht);

/* ---- equi_join ------------------------------------------------------------
 * Inner equi-join producing (left_idx, right_idx) pairs, equi_join<T>
 * (dataframe/join.hpp:304-326, VE body; shrink_missed join.cc:10-61;
 * multi_equi_join join.hpp:198-302).  AUTO = the reference's choice (hash
 * join when the right keys are unique, output in left order; else
 * sort-merge, output in (key, left, right) order).  Outputs hold `capacity`
 * pairs; *nout receives the pair count (also when it does not fit, with
 * status FHIP_ERR_TOO_LARGE).  *used_hash (may be NULL) tells which path
 * ran.  The inputs are not modified. */
#define FHIP_JOIN_AUTO 0
#define FHIP_JOIN_HASH 1
#define FHIP_JOIN_SORTMERGE 2
/* AUTO for a right side the caller expects unique (fhip_learned_join_strategy):
 * the sampled duplicate test is skipped and the table build decides, so the
 * result is AUTO's. */
#define FHIP_JOIN_AUTO_UNIQUE 3
int fhip_dev_equi_join(int ktype, const void* left, const size_t* left_idx,
                       size_t nl, const void* right, const size_t* right_idx,
                       size_t nr, int strategy, size_t* left_out,
                       size_t* right_out, size_t capacity, size_t* nout,
                       int* used_hash, void* stream);

/* Outer equi-join (LEFT outer), outer_equi_join<T> (dataframe/join.hpp:
 * 328-379, VE body): the matched pairs as fhip_dev_equi_join, plus
 * only_left, the left row ids without a match (the rows whose right columns
 * become NULL, hash_joined_dftable, dftable.hpp:874-888).  Hash path:
 * only_left in left order (left_idx[missed], :340-350); sort-merge path:
 * in stable key order of LEFT (:352-378).  only_left holds nl entries;
 * *n_only_left receives the count. */
int fhip_dev_outer_equi_join(int ktype, const void* left,
                             const size_t* left_idx, size_t nl,
                             const void* right, const size_t* right_idx,
                             size_t nr, int strategy, size_t* left_out,
                             size_t* right_out, size_t capacity, size_t* nout,
                             size_t* only_left, size_t* n_only_left,
                             int* used_hash, void* stream);

/* star_join on one rank (dftable.cc:295-333 -> star_join_eq,
 * dfcolumn_join.cc:561-605 -> unique_equi_join2, join.hpp:103-117): the
 * left table (rows left_idx, or 0..nl-1 when NULL) is joined to ndims
 * dimension tables at once, dimension d on left_cols[d] = right_cols[d]
 * (type ktypes[d], nr[d] right rows with ids right_idx[d]).  Every right
 * column must hold unique keys ("star_join: right column is not unique",
 * FHIP_ERR_ARG).  The left rows found in every dimension come out in left
 * order: left_out (nl entries of capacity) and right_out[d] (nl entries
 * each); *nout receives the row count.  At most 8 dimensions; every
 * left_idx entry must be below the left columns' length. */
int fhip_dev_star_join(int ndims, const int* ktypes,
                       const void* const* left_cols, const size_t* left_idx,
                       size_t nl, const void* const* right_cols,
                       const size_t* const* right_idx, const size_t* nr,
                       size_t* left_out, size_t* const* right_out,
                       size_t* nout, void* stream);

/* ---- materialisation, filter and sort of dftable columns -------------------
 * FHIP_NULL_INDEX in a row index selects NULL (an outer join's missing
 * right row).  NULL values are numeric_limits<T>::max(), the reference's
 * NULL value (dfcolumn_misc.cc:1052-1078, append_nulls). */
#define FHIP_NULL_INDEX ((size_t)-1)

/* dst[i] = src[idx[i]], or numeric_limits<T>::max() where idx[i] ==
 * FHIP_NULL_INDEX: global_extract on one rank (dfcolumn_misc.cc:714-735 and
 * the store step :743-761) followed by append_nulls (:1052-1078) when the
 * outer rows' indices are FHIP_NULL_INDEX.  An index >= src_n is an
 * FHIP_ERR_ARG error. */
int fhip_dev_gather(int type, const void* src, size_t src_n,
                    const size_t* idx, size_t n, void* dst, void* stream);
/* The null list of a gathered column: ascending positions i with idx[i] ==
 * FHIP_NULL_INDEX or idx[i] in src_nulls (the source column's null rows):
 * global_extract_null_helper2 (dfcolumn.cc:223-233) plus the positions
 * append_nulls adds.  out holds n entries. */
int fhip_dev_gather_nulls(const size_t* idx, size_t n, size_t src_n,
                          const size_t* src_nulls, size_t nnull, size_t* out,
                          size_t* nout, void* stream);

/* comparison codes of filter (dataframe/dfcolumn_filter_{eq,lt,le,gt,ge}.cc
 * predicates; eq and neq live in dfcolumn_filter_eq.cc) */
#define FHIP_CMP_EQ 0
#define FHIP_CMP_NEQ 1
#define FHIP_CMP_LT 2
#define FHIP_CMP_LE 3
#define FHIP_CMP_GT 4
#define FHIP_CMP_GE 5

/* filter_*_immed on one rank: the ascending row positions i < n where
 * col[i] <op> scalar holds, compared with C++'s usual arithmetic
 * conversions between the column type ctype and the scalar type stype
 * (find_condition, core/find_condition.hpp:23-152, with the templated
 * predicates of dfcolumn_filter_eq.cc:24-35), minus the column's null rows
 * (set_difference with nulls, :100-101).  scalar is a host pointer to one
 * value of stype.  out holds n entries. */
int fhip_dev_filter(int ctype, const void* col, size_t n, int op, int stype,
                    const void* scalar, const size_t* nulls, size_t nnull,
                    size_t* out, size_t* nout, void* stream);
/* filter_eq / neq / lt / le / gt / ge between two columns of one table
 * (find_condition_pair, core/find_condition.hpp:412-546, dfcolumn_filter_eq.cc:
 * 12-21 and :60-102, :160-214).  Null rows
 * of a are removed; those of b too except for eq (the reference skips them
 * there: NULL never equals a non-NULL value). */
int fhip_dev_filter_pair(int atype, const void* a, int btype, const void* b,
                         size_t n, int op, const size_t* a_nulls,
                         size_t na_null, const size_t* b_nulls, size_t nb_null,
                         size_t* out, size_t* nout, void* stream);

/* dftable::sort / sort_desc on one rank (dftable.cc:96-122 ->
 * dfcolumn_sort.cc:69-88 and :141-180): the stable order of the column,
 * ascending (radix_sort) or descending (radix_sort_desc).  idx_in NULL:
 * the rows 0..n-1 of col (n == col_n), the dftable case.  idx_in given: the
 * current order of a sorted_dftable (sorted_dftable::sort, dftable.cc:
 * 2778-2800 -> sort_with_idx, dfcolumn_sort.cc:299-340); the keys are
 * col[idx_in[i]], and sorting again by another column is the reference's
 * multi-column sort ("sort c2, then c1").  Outputs: sorted_out (n keys) and
 * idx_out (n row ids); idx_out may equal idx_in.  Every idx_in entry must be
 * below col_n (the caller's contract, as for fhip_dev_group_by's idx). */
int fhip_dev_sort_index(int ktype, const void* col, size_t col_n,
                        const size_t* idx_in, size_t n, int descending,
                        void* sorted_out, size_t* idx_out, void* stream);

/* dftable::group_by over several key columns on one rank (dftable.cc:
 * 367-487 -> dfcolumn_groupby.cc:122-234): idx (n row ids into every
 * column, permuted in place) is stably sorted by columns ncols-1 .. 1 and
 * then 0 (multi_group_by_sort / _sort_split), so the groups come out in
 * lexicographic key order; split receives the union of every column's run
 * boundaries (merge_split = set_union, dftable.cc:336-339), n + 1 entries
 * of capacity; keys_out[c] (n entries of capacity each) the key of column
 * c of every group (multi_group_by_extract).  The result feeds
 * fhip_dev_grouped_agg unchanged.  Every idx entry must be below the
 * columns' length (the caller's contract, as for fhip_dev_group_by). */
int fhip_dev_multi_group_by(int ncols, const int* ktypes,
                            const void* const* cols, size_t* idx, size_t n,
                            size_t* split, size_t* nsplit,
                            void* const* keys_out, void* stream);

/* ---- host-pointer forms (std::vector storage) ------------------------------
 * Same semantics as the dev_ forms above; inputs are copied to the device
 * and results back.  col_n is the length of the column that idx / gidx
 * index into. */
int fhip_ht_build(int ktype, const void* keys, const uint64_t* vals, size_t n,
                  int check_unique, int* is_unique_ok, fhip_ht** ht);
int fhip_ht_lookup(const fhip_ht* ht, const void* keys, size_t n,
                   uint64_t* vals_out, size_t* miss_out, size_t* miss_n);
int fhip_ht_check_existence(const fhip_ht* ht, const void* keys, size_t n,
                            int* out);
int fhip_ht_all_keys(const fhip_ht* ht, void* out, size_t* out_n);
int fhip_equi_join(int ktype, const void* left, const size_t* left_idx,
                   size_t nl, const void* right, const size_t* right_idx,
                   size_t nr, int strategy, size_t* left_out, size_t* right_out,
                   size_t capacity, size_t* nout, int* used_hash);
/* fhip_equi_join whose inputs are left as the reference's VE equi_join
 * leaves them (dataframe/join.hpp:304-326): when the sort-merge path runs
 * (*used_hash == 0), left / left_idx and right / right_idx come back
 * stable-sorted by key (join.hpp:320-322), from the device copies the join
 * sorted, with no second sort; the hash path leaves them unchanged. */
int fhip_equi_join_inplace(int ktype, void* left, size_t* left_idx, size_t nl,
                           void* right, size_t* right_idx, size_t nr,
                           int strategy, size_t* left_out, size_t* right_out,
                           size_t capacity, size_t* nout, int* used_hash);
int fhip_group_by(int ktype, const void* col, size_t col_n, size_t* idx,
                  size_t n, size_t* split, size_t* nsplit, void* keys_out);
int fhip_grouped_agg(int op, int vtype, const void* col, size_t col_n,
                     const size_t* grouped_idx, size_t n, const size_t* split,
                     size_t nsplit, const size_t* nulls, size_t nnull,
                     void* out);

/* ---- learned CDF model and learned partition sort -------------------------
 * No counterpart in the reference.  The output of fhip_dev_learned_sort is
 * that of radix_sort(K*, size_t) (core/radix_sort.hpp:221-242) on the same
 * keys (64-bit integer keys, key only).
 *  - fhip_learned_train: FP32 MLP 1->64->64->1 trained on the GPU (Adam,
 *    `steps` full-batch steps, learning rate lr) on (key, rank) samples of an
 *    already sorted device array (every n/samples-th key), then distilled
 *    into 1024 coarse splitter keys (the network routes keys to coarse
 *    buckets) over 65536 leaves at the sample quantiles.
 *  - fhip_dev_learned_sort: sorts in place; *fallback_keys gets how many
 *    keys went through the plain radix sort because their leaf overflowed
 *    (n when the whole input fell back); *max_coarse_bucket the largest
 *    coarse bucket (load balance of the split stage).  coarse_mode selects
 *    where the coarse splitters come from: FHIP_COARSE_NETWORK (the
 *    trained network), or, for comparison, the same pipeline without it:
 *    FHIP_COARSE_EQUAL_WIDTH (equal-width ranges of the key space) or
 *    FHIP_COARSE_QUANTILE (the training sample's quantiles b / 1024,
 *    snapped to leaf splitters).
 *  - fhip_learned_cardinality: expected distinct keys among n rows, from
 *    the model: the training column's own count when its keys were all
 *    distinct (n), else n independent draws from the model's CDF.
 *  - fhip_learned_table_hint: that estimate as the table_hint of
 *    fhip_dev_group_by_agg (it sizes the partitioned hash tables).
 *  - fhip_learned_join_strategy: FHIP_JOIN_AUTO_UNIQUE when the model
 *    predicts n_build distinct keys, else FHIP_JOIN_AUTO; the join's output
 *    is the same either way.
 *  - fhip_learned_heavy_keys: how many heavy keys the model kept (the
 *    longest runs of its sorted training column, at most 4096) and the
 *    fraction of the training rows they hold.
 *  - fhip_dev_group_by_agg_model: fhip_dev_group_by_agg with the model of
 *    the key column (trained on an earlier input) in place of the table
 *    hint: its cardinality estimate sizes the partitioned tables and, for a
 *    column of the model's key type, its heavy keys fill the heavy-key
 *    table, so no sample of the input is counted first.  The result is that
 *    of fhip_dev_group_by_agg; a model of another distribution only makes
 *    the step slower.
 *  - fhip_dev_mlp_*: the network's kernels (gradient of the mean squared
 *    error, forward pass, Adam step) on a parameter vector of
 *    fhip_mlp_nparam() floats. */
#define FHIP_COARSE_NETWORK 0
#define FHIP_COARSE_EQUAL_WIDTH 1
#define FHIP_COARSE_QUANTILE 2
typedef struct fhip_cdf_model fhip_cdf_model;
int fhip_learned_train(int ktype, const void* sorted_keys, size_t n,
                       size_t samples, int steps, float lr,
                       fhip_cdf_model** model, void* stream);
int fhip_dev_learned_sort(const fhip_cdf_model* m, void* keys, size_t n,
                          int coarse_mode, size_t* fallback_keys,
                          size_t* fallback_buckets, size_t* max_coarse_bucket,
                          void* stream);
int fhip_learned_cardinality(const fhip_cdf_model* m, size_t n,
                             double* distinct);
int fhip_learned_table_hint(const fhip_cdf_model* m, size_t n,
                            size_t* table_hint);
int fhip_learned_join_strategy(const fhip_cdf_model* m, size_t n_build,
                               int* strategy);
int fhip_learned_heavy_keys(const fhip_cdf_model* m, size_t* n_keys,
                            double* row_frac);
int fhip_dev_group_by_agg_model(int ktype, const void* keys, size_t n,
                                const fhip_agg* aggs, int naggs,
                                void* keys_out, size_t capacity,
                                size_t* ngroups, const fhip_cdf_model* model,
                                void* stream);
int fhip_learned_model_info(const fhip_cdf_model* m, double* final_loss,
                            double* cdf_knots, uint64_t* coarse_splitters);
int fhip_learned_model_params(const fhip_cdf_model* m, float* params);
int fhip_learned_model_destroy(fhip_cdf_model* m);
int fhip_mlp_nparam(void);
int fhip_dev_mlp_grad(const float* params, const float* x, const float* t,
                      size_t S, float* grad, float* loss, void* stream);
int fhip_dev_mlp_forward(const float* params, const float* x, size_t S,
                         float* y, void* stream);
int fhip_dev_mlp_adam(float* params, float* grad, float* m, float* v,
                      int step, float lr, void* stream);

/* ---- batched segments --------------------------------------------------------
 * No counterpart in the reference: runs, for each of nseg segments
 * (rows [seg_off[s], seg_off[s+1]) of keys/vals, seg_off on the device), the
 * reference's radix_sort(key, val) (core/radix_sort.hpp:197-219) and the
 * per-group sum and count of dfcolumn_grouped_agg.cc (sequential fold,
 * :153-164), all segments in one HIP graph launch.  Keys must lie in
 * [0, 2^key_bits).  Outputs: keys_out/vals_out (n, sorted per segment,
 * stable), per group gkeys/gsum/gcount (capacity entries; groups of segment
 * s are [gseg_off[s], gseg_off[s+1])), *ngroups.  use_graph = 1 captures
 * the launch sequence once per shape and replays it. */
int fhip_dev_batch_sort_groupby(const int64_t* keys, const double* vals,
                                size_t n, const size_t* seg_off, size_t nseg,
                                int key_bits, int64_t* keys_out,
                                double* vals_out, int64_t* gkeys, double* gsum,
                                uint64_t* gcount, size_t* gseg_off,
                                size_t capacity, size_t* ngroups,
                                int use_graph, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FROVEDIS_HIP_H */
