// CPU oracle for the frovedis per-rank dataframe execution core.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain C++ on the host,
// the algorithms of the reference's x86 code path so that the HIP library
// can be checked against it.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load it; the product library never does.
//
// Every function cites the reference file:line it follows (paths relative to
// the frovedis source tree).  How each function is pinned (closed forms,
// tutorial known answers, brute force) is stated next to it; the pinning
// tests live in tests/test_oracle.py.
//
// Type codes shared with the HIP library: 0=int32 1=uint32 2=int64 3=uint64
// 4=float 5=double.

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <limits>
#include <stdexcept>
#include <unordered_map>
#include <vector>

namespace {

enum { T_I32 = 0, T_U32 = 1, T_I64 = 2, T_U64 = 3, T_F32 = 4, T_F64 = 5 };

// ---------------------------------------------------------------------------
// radix_sort: src/frovedis/core/radix_sort.hpp:197-242 (dispatch, flips,
// insertion threshold 64 on x86 :13-18) and radix_sort.incl:7-240 (one stable
// counting pass per byte of the key, least significant first).
// Pinned by: closed form (numpy stable argsort) for integer keys; for float
// keys by the quirks read from radix_sort.hpp:101-118,244-266 (tests).
// ---------------------------------------------------------------------------

constexpr size_t kInsertionThr = 64;  // radix_sort.hpp:17 (x86)

// insertion sort with the reference's comparison (`>`, or `<` for desc),
// radix_sort.hpp:31-99
template <class K>
void insertion(K* k, uint8_t* v, int vb, size_t n, bool desc) {
  for (size_t i = 1; i < n; i++) {
    K tmp = k[i];
    uint8_t vtmp[8];
    if (vb) std::memcpy(vtmp, v + i * vb, vb);
    bool out_of_order = desc ? (k[i - 1] < tmp) : (k[i - 1] > tmp);
    if (!out_of_order) continue;
    size_t j = i;
    do {
      k[j] = k[j - 1];
      if (vb) std::memcpy(v + j * vb, v + (j - 1) * vb, vb);
      j--;
    } while (j > 0 && (desc ? (k[j - 1] < tmp) : (k[j - 1] > tmp)));
    k[j] = tmp;
    if (vb) std::memcpy(v + j * vb, vtmp, vb);
  }
}

// One LSD pass per byte for bytes [0, nbytes); radix_sort.incl:43-228.
// Descending uses bucket 255-digit (radix_sort.incl:242-480).
template <class U>
void lsd(U* k, uint8_t* v, int vb, size_t n, size_t nbytes, bool desc) {
  std::vector<U> kt(n);
  std::vector<uint8_t> vt(vb ? n * vb : 0);
  U* ks = k; U* kd = kt.data();
  uint8_t* vs = v; uint8_t* vd = vt.data();
  for (size_t d = 0; d < nbytes; d++) {
    size_t cnt[257] = {0};
    int sh = int(d * 8);
    for (size_t i = 0; i < n; i++) {
      unsigned b = unsigned((ks[i] >> sh) & 0xFF);
      if (desc) b = 255 - b;
      cnt[b + 1]++;
    }
    for (int b = 0; b < 256; b++) cnt[b + 1] += cnt[b];
    for (size_t i = 0; i < n; i++) {
      unsigned b = unsigned((ks[i] >> sh) & 0xFF);
      if (desc) b = 255 - b;
      size_t to = cnt[b]++;
      kd[to] = ks[i];
      if (vb) std::memcpy(vd + to * vb, vs + i * vb, vb);
    }
    std::swap(ks, kd);
    std::swap(vs, vd);
  }
  if (ks != k) {  // odd pass count: copy back (radix_sort.incl:230-239)
    std::memcpy(k, ks, n * sizeof(U));
    if (vb) std::memcpy(v, vs, n * vb);
  }
}

// get_max_key_size: radix_sort.hpp:135-153 (number of bytes of the maximum,
// maximum initialised to 0)
template <class K>
size_t max_key_size(const K* k, size_t n) {
  K mx = 0;
  for (size_t i = 0; i < n; i++) if (k[i] > mx) mx = k[i];
  size_t ks = sizeof(K);
  unsigned long long tmp = 0xFF;
  for (size_t i = 1; i < ks + 1; i++) {
    if ((unsigned long long)mx <= tmp) return i;
    tmp = (tmp << 8) | 0xFF;
  }
  return ks;
}

template <class K>
bool contain_negative(const K* k, size_t n) {  // radix_sort.hpp:101-133
  if (!std::numeric_limits<K>::is_signed && std::numeric_limits<K>::is_integer)
    return false;
  for (size_t i = 0; i < n; i++) if (k[i] < 0) return true;
  return false;
}

template <class K, class U>
void radix_sort_int(K* k, uint8_t* v, int vb, size_t n, bool positive_only,
                    bool desc) {
  if (n < kInsertionThr) { insertion(k, v, vb, n, desc); return; }
  U* u = reinterpret_cast<U*>(k);
  const U top = U(1) << (sizeof(U) * 8 - 1);
  if (positive_only) {
    lsd(u, v, vb, n, max_key_size(k, n), desc);
    return;
  }
  bool neg = contain_negative(k, n);
  if (neg) {
    for (size_t i = 0; i < n; i++) u[i] ^= top;          // flip_top :155-166
    lsd(u, v, vb, n, sizeof(U), desc);
    for (size_t i = 0; i < n; i++) u[i] ^= top;
  } else {
    lsd(u, v, vb, n, max_key_size(k, n), desc);
  }
}

template <class F, class U>
void radix_sort_fp(F* k, uint8_t* v, int vb, size_t n, bool positive_only,
                   bool desc) {
  if (n < kInsertionThr) { insertion(k, v, vb, n, desc); return; }
  U* u = reinterpret_cast<U*>(k);
  const int sh = int(sizeof(U) * 8 - 1);
  const U top = U(1) << sh;
  if (positive_only) {  // radix_sort.hpp:250-253: raw bits
    lsd(u, v, vb, n, max_key_size(u, n), desc);
    return;
  }
  bool neg = contain_negative(k, n);
  if (neg) {
    // float_flip / double_flip, radix_sort.hpp:168-195
    for (size_t i = 0; i < n; i++) u[i] ^= (U(0) - (u[i] >> sh)) | top;
    lsd(u, v, vb, n, sizeof(U), desc);
    for (size_t i = 0; i < n; i++) u[i] ^= ((u[i] >> sh) - 1) | top;
  } else {
    lsd(u, v, vb, n, max_key_size(u, n), desc);
  }
}

// radix_sort(K*, size_t*, size_t) / radix_sort_desc of a key column with a
// row-id payload, by key type (core/radix_sort.hpp:197-219, :244-340,
// :342-487).
template <class K>
void sort_kv(K* k, size_t* v, size_t n, bool desc) {
  uint8_t* vb = reinterpret_cast<uint8_t*>(v);
  if (std::numeric_limits<K>::is_integer) {
    if (sizeof(K) == 4) {
      if (std::numeric_limits<K>::is_signed)
        radix_sort_int<int32_t, uint32_t>(reinterpret_cast<int32_t*>(k), vb, 8, n, false, desc);
      else
        radix_sort_int<uint32_t, uint32_t>(reinterpret_cast<uint32_t*>(k), vb, 8, n, false, desc);
    } else {
      if (std::numeric_limits<K>::is_signed)
        radix_sort_int<int64_t, uint64_t>(reinterpret_cast<int64_t*>(k), vb, 8, n, false, desc);
      else
        radix_sort_int<uint64_t, uint64_t>(reinterpret_cast<uint64_t*>(k), vb, 8, n, false, desc);
    }
  } else if (sizeof(K) == 4) {
    radix_sort_fp<float, uint32_t>(reinterpret_cast<float*>(k), vb, 8, n, false, desc);
  } else {
    radix_sort_fp<double, uint64_t>(reinterpret_cast<double*>(k), vb, 8, n, false, desc);
  }
}

// ---------------------------------------------------------------------------
// set_separate / set_unique / set_is_unique: core/set_operations.hpp:146-197
// (x86 bodies).  Pinned by closed form (run starts of a sorted array).
// ---------------------------------------------------------------------------
template <class T>
size_t separate(const T* k, size_t n, size_t* out) {
  if (n == 0) { out[0] = 0; return 1; }
  size_t c = 0;
  T cur = k[0];
  out[c++] = 0;
  for (size_t i = 1; i < n; i++)
    if (k[i] != cur) { cur = k[i]; out[c++] = i; }
  out[c++] = n;
  return c;
}

template <class T>
int is_unique(const T* k, size_t n) {
  if (n == 0) return 1;
  T cur = k[0];
  for (size_t i = 1; i < n; i++) {
    if (k[i] != cur) cur = k[i];
    else return 0;
  }
  return 1;
}

// lower_bound / upper_bound: core/lower_bound.hpp:13-101,
// core/upper_bound.hpp:13-101.  The reference halves [low, high) with
// mid = (low + high) / 2 and its own comparisons: lower goes left when
// `x <= sorted[mid]`, upper goes right when `x >= sorted[mid]`.  On ordered
// inputs that is std::lower_bound / std::upper_bound; a NaN query goes
// right in lower_bound (result m) and left in upper_bound (result 0).
// Pinned by closed form (numpy searchsorted) and those NaN special cases.
template <class T>
size_t ref_lower(const T* s, size_t m, T x) {
  size_t lo = 0, hi = m;
  while (lo < hi) {
    size_t mid = (lo + hi) / 2;
    if (x <= s[mid]) hi = mid; else lo = mid + 1;
  }
  return lo;
}
template <class T>
void lower_bound_t(const T* s, size_t m, const T* q, size_t nq, size_t* out) {
  for (size_t i = 0; i < nq; i++) out[i] = ref_lower(s, m, q[i]);
}
template <class T>
void upper_bound_t(const T* s, size_t m, const T* q, size_t nq, size_t* out) {
  for (size_t i = 0; i < nq; i++) {
    size_t lo = 0, hi = m;
    while (lo < hi) {
      size_t mid = (lo + hi) / 2;
      if (q[i] >= s[mid]) lo = mid + 1; else hi = mid;
    }
    out[i] = lo;
  }
}

// ---------------------------------------------------------------------------
// unique_hashtable<K, size_t>: dataframe/hashtable.hpp.
//  - table size: smallest listed prime >= 3n (:23, :115-152, :1174-1179)
//  - home slot: unsigned key bits % table_size (:1190-1212)
//  - the last row (in index order) to hash to a slot owns it (:1197-1200)
//  - rows whose key differs from the slot owner's key form the conflict list,
//    in row order (find_miss_* :155-1165), then stable radix-sorted (:1234)
//  - uniqueness: a row equal to the owner's key that is not the owner, or a
//    duplicate inside the sorted conflict list (:1266-1321)
//  - lookup: slot key, else lower_bound into the conflict list (:1386-1515)
//  - all_keys: owned slots in slot order, then the conflict keys (:1584-1607)
// Pinned by: dictionary semantics and brute force (tests).
// ---------------------------------------------------------------------------
const long long kPrimes[] = {
    8 + 3, 16 + 3, 32 + 5, 64 + 3, 128 + 3, 256 + 27, 512 + 9, 1024 + 9,
    2048 + 5, 4096 + 3, 8192 + 27, 16384 + 43, 32768 + 3, 65536 + 45,
    131072 + 29, 262144 + 3, 524288 + 21, 1048576 + 7, 2097152 + 17,
    4194304 + 15, 8388608 + 9, 16777216 + 43, 33554432 + 35, 67108864 + 15,
    134217728 + 29, 268435456 + 3, 536870912 + 11, 1073741824 + 85,
    2147483648LL + 11, 4294967296LL + 15, 8589934592LL + 17,
    17179869184LL + 25, 34359738368LL + 53};

struct HT {
  int kbytes;
  size_t table_size;
  std::vector<uint64_t> key;      // key bits per slot
  std::vector<uint64_t> val;
  std::vector<int> filled;
  std::vector<uint64_t> ckey;     // conflict keys (bits), sorted by typed key
  std::vector<uint64_t> cval;
  int ktype;
};

inline uint64_t load_bits(const void* p, int kb, size_t i) {
  if (kb == 4) return reinterpret_cast<const uint32_t*>(p)[i];
  return reinterpret_cast<const uint64_t*>(p)[i];
}

template <class K>
K as_typed(uint64_t bits) {
  K k;
  if (sizeof(K) == 4) { uint32_t b = uint32_t(bits); std::memcpy(&k, &b, 4); }
  else std::memcpy(&k, &bits, 8);
  return k;
}

template <class K>
bool typed_eq(uint64_t a, uint64_t b) { return as_typed<K>(a) == as_typed<K>(b); }

size_t table_size_for(size_t n) {
  size_t target = n * 3;
  for (long long p : kPrimes) if ((unsigned long long)p >= target) return size_t(p);
  throw std::runtime_error("unique_hash_table: requested size too large");
}

template <class K>
void ht_build_t(HT* h, const void* keys, const uint64_t* vals, size_t n,
                int check_unique, int* is_unique_ok) {
  h->table_size = table_size_for(n);
  size_t ts = h->table_size;
  h->key.assign(ts, 0);
  h->val.assign(ts, 0);
  h->filled.assign(ts, 0);
  std::vector<size_t> owner(ts, std::numeric_limits<size_t>::max());
  std::vector<size_t> hash(n);
  for (size_t i = 0; i < n; i++) {
    uint64_t b = load_bits(keys, h->kbytes, i);
    size_t s = size_t(b % ts);
    h->key[s] = b;
    h->filled[s] = 1;
    owner[s] = i;
    hash[i] = s;
  }
  int uniq = 1;
  std::vector<size_t> missed;
  for (size_t i = 0; i < n; i++) {
    uint64_t b = load_bits(keys, h->kbytes, i);
    size_t s = hash[i];
    if (!typed_eq<K>(b, h->key[s])) { missed.push_back(i); continue; }
    if (owner[s] == i) { if (vals) h->val[s] = vals[i]; }
    else {
      // duplicate of the owner's key: only the unique-check ctor records it
      uniq = 0;
      if (!check_unique && vals) h->val[s] = vals[i];  // unspecified in ref
    }
  }
  // conflict list in row order, then radix_sort(conflict_key, conflict_val)
  // (:1234): the reference's own sort rules, so float keys follow its
  // flips (-0.0 after the positives when no conflict key is negative, NaN
  // by bit pattern) and a list under 64 keys its insertion sort
  std::vector<K> ck(missed.size());
  std::vector<size_t> order(missed.size());
  for (size_t i = 0; i < order.size(); i++) {
    ck[i] = as_typed<K>(load_bits(keys, h->kbytes, missed[i]));
    order[i] = i;
  }
  sort_kv<K>(ck.data(), order.data(), ck.size(), false);
  h->ckey.resize(missed.size());
  h->cval.resize(missed.size());
  for (size_t i = 0; i < order.size(); i++) {
    size_t r = missed[order[i]];
    h->ckey[i] = load_bits(keys, h->kbytes, r);
    h->cval[i] = vals ? vals[r] : 0;
  }
  for (size_t i = 1; i < h->ckey.size(); i++)
    if (typed_eq<K>(h->ckey[i], h->ckey[i - 1])) uniq = 0;
  if (is_unique_ok) *is_unique_ok = uniq;
}

// lower_bound(conflict_key, missed_key) (:1430, core/lower_bound.hpp:13-101)
template <class K>
size_t conflict_lb(const HT* h, uint64_t b) {
  size_t lo = 0, hi = h->ckey.size();
  const K x = as_typed<K>(b);
  while (lo < hi) {
    size_t mid = (lo + hi) / 2;
    if (x <= as_typed<K>(h->ckey[mid])) hi = mid; else lo = mid + 1;
  }
  return lo;
}

// lookup with miss list (hashtable.hpp:1442-1515): misses are returned as the
// sorted union of "clear" misses (empty slot) and conflict-list misses; the
// value of a missed row stays 0 (value-initialised output, :1447).
template <class K>
size_t ht_lookup_t(const HT* h, const void* keys, size_t n, uint64_t* out,
                   size_t* miss) {
  size_t nm = 0;
  for (size_t i = 0; i < n; i++) {
    uint64_t b = load_bits(keys, h->kbytes, i);
    size_t s = size_t(b % h->table_size);
    out[i] = 0;
    if (!h->filled[s]) { miss[nm++] = i; continue; }
    if (typed_eq<K>(h->key[s], b)) { out[i] = h->val[s]; continue; }
    size_t p = conflict_lb<K>(h, b);
    if (p < h->ckey.size() && typed_eq<K>(h->ckey[p], b)) out[i] = h->cval[p];
    else miss[nm++] = i;
  }
  return nm;
}

template <class K>
void ht_exist_t(const HT* h, const void* keys, size_t n, int* out) {
  for (size_t i = 0; i < n; i++) {
    uint64_t b = load_bits(keys, h->kbytes, i);
    size_t s = size_t(b % h->table_size);
    if (!h->filled[s]) { out[i] = 0; continue; }
    if (typed_eq<K>(h->key[s], b)) { out[i] = 1; continue; }
    size_t p = conflict_lb<K>(h, b);
    out[i] = (p < h->ckey.size() && typed_eq<K>(h->ckey[p], b)) ? 1 : 0;
  }
}

// ---------------------------------------------------------------------------
// equi_join: dataframe/join.hpp.
//  algo 0 (VE body, :304-326): unique_hashtable on RIGHT; if unique, probe
//    LEFT in order and drop misses (shrink_missed, join.cc:10-61); otherwise
//    stable-sort both sides and emit, for every left row in sorted order, its
//    matching right rows in sorted order (multi_equi_join :198-302).
//  algo 1 (x86 body, :25-45): std::unordered_multimap on RIGHT, probe LEFT in
//    order; the order inside a key follows libstdc++ and is only compared
//    after canonical sorting.
//  algo 2: the sort-merge branch of algo 0 without the table build.
// With only_left, the LEFT outer join outer_equi_join (VE :328-379, x86
// :73-99): the left row ids without a match -- in left order on the hash
// paths, in stable key order of LEFT on the sort-merge path.
// Pinned by: brute force pair sets, the tutorial join and outer join
// (doc/tutorial/tutorial.md:2362-2375, :2456-2466), and for only_left the
// closed form "left rows whose key is absent from RIGHT" (tests).
// ---------------------------------------------------------------------------
template <class K>
void equi_join_t(const void* left, const size_t* lidx, size_t nl,
                 const void* right, const size_t* ridx, size_t nr, int algo,
                 std::vector<size_t>& lo, std::vector<size_t>& ro,
                 int* used_hash, std::vector<size_t>* only_left) {
  const K* l = reinterpret_cast<const K*>(left);
  const K* r = reinterpret_cast<const K*>(right);
  lo.clear(); ro.clear();
  if (only_left) only_left->clear();
  if (algo == 1) {
    std::unordered_multimap<K, size_t> m;
    for (size_t i = 0; i < nr; i++) m.insert(std::make_pair(r[i], ridx[i]));
    for (size_t i = 0; i < nl; i++) {
      auto range = m.equal_range(l[i]);
      if (range.first == range.second && only_left) only_left->push_back(lidx[i]);
      for (auto it = range.first; it != range.second; ++it) {
        lo.push_back(lidx[i]);
        ro.push_back(it->second);
      }
    }
    if (used_hash) *used_hash = 1;
    return;
  }
  int uniq = 0;
  if (algo != 2) {
    HT h;
    h.kbytes = sizeof(K);
    uniq = 1;
    std::vector<uint64_t> rv(ridx, ridx + nr);
    ht_build_t<K>(&h, right, rv.data(), nr, 1, &uniq);
    if (uniq) {
      std::vector<uint64_t> got(nl);
      std::vector<size_t> miss(nl);
      size_t nm = ht_lookup_t<K>(&h, left, nl, got.data(), miss.data());
      size_t mi = 0;
      for (size_t i = 0; i < nl; i++) {
        if (mi < nm && miss[mi] == i) { mi++; continue; }
        lo.push_back(lidx[i]);
        ro.push_back(size_t(got[i]));
      }
      // outer: left_idx[missed] (join.hpp:340-350)
      if (only_left)
        for (size_t t = 0; t < nm; t++) only_left->push_back(lidx[miss[t]]);
      if (used_hash) *used_hash = 1;
      return;
    }
  }
  // sort-merge: radix_sort(right, right_idx); radix_sort(left, left_idx)
  // (join.hpp:320-322), then every sorted left row against its run of equal
  // right keys (multi_equi_join :198-302); outer: the sorted left rows whose
  // run is empty (:352-378)
  std::vector<K> lk(l, l + nl), rk(r, r + nr);
  std::vector<size_t> li(lidx, lidx + nl), ri(ridx, ridx + nr);
  sort_kv(rk.data(), ri.data(), nr, false);
  sort_kv(lk.data(), li.data(), nl, false);
  size_t j = 0;
  for (size_t ii = 0; ii < nl; ii++) {
    K key = lk[ii];
    while (j < nr && rk[j] < key) j++;
    size_t t = j;
    for (; t < nr && rk[t] == key; t++) {
      lo.push_back(li[ii]);
      ro.push_back(ri[t]);
    }
    if (t == j && only_left) only_left->push_back(li[ii]);
  }
  if (used_hash) *used_hash = 0;
}

// ---------------------------------------------------------------------------
// group_by on one rank: dataframe/dfcolumn_groupby.cc:16-105.  With one node,
// split_by_hash / alltoall_exchange / the second sort are identities and
// merge_map[0] = iota, so the per-rank result is: newval[i] = val[idx[i]];
// radix_sort(newval, idx); split = set_separate(newval); group keys =
// newval[split[i]].  The last group is the NULL group when its key equals
// numeric_limits<T>::max() (:98-102).
// ---------------------------------------------------------------------------
template <class K>
size_t group_by_t(const void* col, size_t* idx, size_t n, size_t* split,
                  void* keys_out) {
  const K* c = reinterpret_cast<const K*>(col);
  std::vector<K> nv(n);
  for (size_t i = 0; i < n; i++) nv[i] = c[idx[i]];
  sort_kv(nv.data(), idx, n, false);
  size_t ns = separate(nv.data(), n, split);
  K* ko = reinterpret_cast<K*>(keys_out);
  for (size_t i = 0; i + 1 < ns; i++) ko[i] = nv[split[i]];
  return ns;
}

// ---------------------------------------------------------------------------
// Grouped aggregation on one rank: dataframe/dfcolumn_grouped_agg.cc (x86
// #else branches).  With one node the exchange is an identity and the merge
// is newval[j] = T() + partial[j] (0 + x; :213-232), so:
//   sum  (:105-190):  nulls read as 0; sequential left fold from 0 in
//                     grouped (stable sorted) row order; result 0 + fold.
//   count(:719-813):  split[i+1]-split[i] without nulls; else a sum of ones
//                     with nulls read as 0.
//   max  (:366-452):  nulls read as lowest(); m = lowest(); if (m < v) m = v.
//   min  (:513-717):  nulls read as max(); m = max(); if (m > v) m = v.
//   avg  (:839-889):  double(sum)/double(count); groups whose count is 0 get
//                     numeric_limits<T>::max().
// op codes: 0=sum 1=count 2=avg 3=max 4=min.  Outputs: sum/max/min in T,
// count as uint64, avg as double.  Pinned by: tutorial known answers and
// closed forms (tests); the FP fold order is the restatement itself.
// ---------------------------------------------------------------------------
template <class T>
void grouped_agg_t(int op, const void* colp, const size_t* gidx, size_t n,
                   const size_t* split, size_t ns, const size_t* nulls,
                   size_t nn, void* out) {
  const T* col = reinterpret_cast<const T*>(colp);
  std::vector<uint8_t> isnull(nn ? n : 0, 0);
  for (size_t i = 0; i < nn; i++) isnull[nulls[i]] = 1;
  size_t G = ns ? ns - 1 : 0;
  auto val_at = [&](size_t row, T nullval) -> T {
    return (nn && isnull[row]) ? nullval : col[row];
  };
  if (op == 0 || op == 2) {
    std::vector<T> sums(G);
    std::vector<uint64_t> cnt(G);
    for (size_t g = 0; g < G; g++) {
      T total = 0;
      uint64_t c = 0;
      for (size_t j = split[g]; j < split[g + 1]; j++) {
        total += val_at(gidx[j], T(0));
        c += (nn && isnull[gidx[j]]) ? 0 : 1;
      }
      T merged = T();
      merged += total;
      sums[g] = merged;
      cnt[g] = c;
    }
    if (op == 0) {
      T* o = reinterpret_cast<T*>(out);
      for (size_t g = 0; g < G; g++)
        o[g] = (nn && cnt[g] == 0) ? std::numeric_limits<T>::max() : sums[g];
    } else {
      double* o = reinterpret_cast<double*>(out);
      for (size_t g = 0; g < G; g++)
        o[g] = (cnt[g] == 0 && nn) ? double(std::numeric_limits<T>::max())
                                   : double(sums[g]) / double(cnt[g]);
    }
  } else if (op == 1) {
    uint64_t* o = reinterpret_cast<uint64_t*>(out);
    for (size_t g = 0; g < G; g++) {
      uint64_t c = 0;
      for (size_t j = split[g]; j < split[g + 1]; j++)
        c += (nn && isnull[gidx[j]]) ? 0 : 1;
      o[g] = c;
    }
  } else if (op == 3 || op == 4) {
    // typed max (:454-510): the merge starts from lowest() and keeps the
    // larger value; with nulls in the column, a result equal to lowest()
    // is NULL and reads max().  min is the mirror image (:513-717).
    T* o = reinterpret_cast<T*>(out);
    const T lo = std::numeric_limits<T>::lowest(), hi = std::numeric_limits<T>::max();
    for (size_t g = 0; g < G; g++) {
      T m = (op == 3) ? lo : hi;
      for (size_t j = split[g]; j < split[g + 1]; j++) {
        T v = val_at(gidx[j], op == 3 ? lo : hi);
        if (op == 3) { if (m < v) m = v; } else { if (m > v) m = v; }
      }
      T merged = (op == 3) ? lo : hi;
      if (op == 3) { if (merged < m) merged = m; } else { if (merged > m) merged = m; }
      if (op == 3 && nn && merged == lo) merged = hi;
      o[g] = merged;
    }
  }
}

template <class F>
void dispatch_key(int t, F&& f) {
  switch (t) {
    case T_I32: f(int32_t()); break;
    case T_U32: f(uint32_t()); break;
    case T_I64: f(int64_t()); break;
    case T_U64: f(uint64_t()); break;
    case T_F32: f(float()); break;
    case T_F64: f(double()); break;
    default: throw std::runtime_error("bad type code");
  }
}

// ---------------------------------------------------------------------------
// star_join on one rank: dftable.cc:295-333.  ret_left_idx starts as the
// left rows; for every dimension in order, unique_equi_join2
// (join.hpp:103-117: unique_hashtable on the right keys with the right ids,
// "star_join: right column is not unique" otherwise, lookup of the left keys
// at ret_left_idx) gives the right ids and the missed positions, and
// shrink_missed_inplace removes the missed rows from ret_left_idx, this
// dimension's right ids and every earlier dimension's.
// Pinned by: brute force (the left rows whose key is present in every
// dimension, in left order, with the position of that key) (tests).
// ---------------------------------------------------------------------------
template <class K>
void star_dim_t(const void* left_col, const void* right_col, const size_t* ridx, size_t nr,
                std::vector<size_t>& lrows, std::vector<std::vector<size_t>>& rids) {
  const K* l = reinterpret_cast<const K*>(left_col);
  HT h;
  h.kbytes = sizeof(K);
  int uniq = 1;
  std::vector<uint64_t> rv(ridx, ridx + nr);
  ht_build_t<K>(&h, right_col, rv.data(), nr, 1, &uniq);
  if (!uniq) throw std::runtime_error("star_join: right column is not unique");
  std::vector<K> q(lrows.size());
  for (size_t i = 0; i < lrows.size(); i++) q[i] = l[lrows[i]];
  std::vector<uint64_t> got(q.size());
  std::vector<size_t> miss(q.size());
  size_t nm = ht_lookup_t<K>(&h, q.data(), q.size(), got.data(), miss.data());
  rids.emplace_back(got.begin(), got.end());
  // shrink_missed_inplace on every output so far
  std::vector<char> drop(lrows.size(), 0);
  for (size_t t = 0; t < nm; t++) drop[miss[t]] = 1;
  auto shrink = [&](std::vector<size_t>& v) {
    size_t c = 0;
    for (size_t i = 0; i < v.size(); i++) if (!drop[i]) v[c++] = v[i];
    v.resize(c);
  };
  shrink(lrows);
  for (auto& r : rids) shrink(r);
}

// ---------------------------------------------------------------------------
// Materialisation of joined / filtered tables on one rank:
//  gather: global_extract (dataframe/dfcolumn_misc.cc:714-761) is
//    ret[i] = val[idx[i]] once the exchange is an identity; append_nulls
//    (:1052-1078) adds numeric_limits<T>::max() rows, here selected by the
//    index ~0 (FHIP_NULL_INDEX).  An index past the column is an error.
//  gather_nulls: positions i whose source row is null
//    (global_extract_null_helper2, dfcolumn.cc:223-233: check_existence of
//    idx[i] in the null set) or whose index is ~0 (the appended rows).
// Pinned by: closed forms (numpy take / isin, tests).
// ---------------------------------------------------------------------------
constexpr size_t kNullIndex = ~size_t(0);

template <class T>
int gather_t(const T* src, size_t src_n, const size_t* idx, size_t n, T* dst) {
  for (size_t i = 0; i < n; i++) {
    if (idx[i] == kNullIndex) dst[i] = std::numeric_limits<T>::max();
    else if (idx[i] < src_n) dst[i] = src[idx[i]];
    else return -1;
  }
  return 0;
}

size_t gather_nulls_t(const size_t* idx, size_t n, const size_t* nulls, size_t nn,
                      size_t* out) {
  std::vector<size_t> s(nulls, nulls + nn);
  std::sort(s.begin(), s.end());
  size_t c = 0;
  for (size_t i = 0; i < n; i++)
    if (idx[i] == kNullIndex || std::binary_search(s.begin(), s.end(), idx[i])) out[c++] = i;
  return c;
}

// ---------------------------------------------------------------------------
// filter on one rank: find_condition / find_condition_pair
// (core/find_condition.hpp:23-152, :412-546) with the predicates of
// dataframe/dfcolumn_filter_{eq,lt,le,gt,ge}.cc -- `a op b` on T and U
// with C++'s usual arithmetic conversions -- then set_difference with the
// sorted null lists: the column's own nulls for every filter, the other
// column's nulls too for neq/lt/le/gt/ge (dfcolumn_filter_eq.cc:60-102,
// :160-214; dfcolumn_filter_lt.cc:37-88).
// op codes: 0 eq, 1 neq, 2 lt, 3 le, 4 gt, 5 ge.
// Pinned by: the tutorial filters (doc/tutorial/src/tut5.3: int c1 against
// double c4, double c2 against 30.0) and numpy comparisons in the common
// type that C++ picks (tests).
// ---------------------------------------------------------------------------
template <class A, class B>
bool cmp_op(A a, B b, int op) {
  switch (op) {
    case 0: return a == b;
    case 1: return a != b;
    case 2: return a < b;
    case 3: return a <= b;
    case 4: return a > b;
    case 5: return a >= b;
    default: throw std::runtime_error("bad comparison code");
  }
}

size_t minus_nulls(std::vector<size_t>& hit, std::vector<size_t> nulls, size_t* out) {
  std::sort(nulls.begin(), nulls.end());
  nulls.erase(std::unique(nulls.begin(), nulls.end()), nulls.end());
  size_t* e = std::set_difference(hit.begin(), hit.end(), nulls.begin(), nulls.end(), out);
  return size_t(e - out);
}

template <class T, class U>
size_t filter_t(const T* col, size_t n, int op, U c, const size_t* nulls, size_t nn,
                size_t* out) {
  std::vector<size_t> hit;
  for (size_t i = 0; i < n; i++) if (cmp_op(col[i], c, op)) hit.push_back(i);
  return minus_nulls(hit, std::vector<size_t>(nulls, nulls + nn), out);
}

template <class T, class U>
size_t filter_pair_t(const T* a, const U* b, size_t n, int op, const size_t* an, size_t na,
                     const size_t* bn, size_t nb, size_t* out) {
  std::vector<size_t> hit;
  for (size_t i = 0; i < n; i++) if (cmp_op(a[i], b[i], op)) hit.push_back(i);
  std::vector<size_t> nulls(an, an + na);
  if (op != 0) nulls.insert(nulls.end(), bn, bn + nb);
  return minus_nulls(hit, nulls, out);
}

// ---------------------------------------------------------------------------
// dftable::sort / sort_desc and sorted_dftable::sort on one rank
// (dataframe/dfcolumn_sort.cc:69-88, :141-180, :299-340): the keys
// val[idx[i]] (idx = iota for a dftable) are sorted with idx as payload by
// radix_sort / radix_sort_desc; with one node the partition, exchange and
// set_multimerge_pair leave that order unchanged.
// Pinned by: numpy stable argsort (closed form) and the tutorial's two-key
// sort (doc/tutorial/tutorial.md:2321-2341, tut5.4/t.csv) (tests).
// ---------------------------------------------------------------------------
template <class K>
void sort_index_t(const K* col, const size_t* idx_in, size_t n, bool desc, K* sorted,
                  size_t* idx_out) {
  for (size_t i = 0; i < n; i++) {
    idx_out[i] = idx_in ? idx_in[i] : i;
    sorted[i] = col[idx_out[i]];
  }
  sort_kv(sorted, idx_out, n, desc);
}

// ---------------------------------------------------------------------------
// dftable::group_by over several columns on one rank (dataframe/dftable.cc:
// 367-487 -> dfcolumn_groupby.cc:122-234): multi_group_by_sort of columns
// size-1 .. 1, multi_group_by_sort_split of column 0 (sort, then
// set_separate), merge_split = set_union with multi_group_by_split of the
// other columns (dftable.cc:336-339), multi_group_by_extract of the keys
// (val[idx[split[g]]]).  With one node the hash exchange and the second
// round of sorts keep that result (the keys are already sorted and unique).
// Pinned by: the tutorial's two-column group_by (doc/tutorial/tutorial.md:
// 2535-2552, tut5.6/t.csv, as a set of rows) and a closed form (Python
// sort by key tuples) (tests).
// ---------------------------------------------------------------------------
// sorts (col[idx], idx) by one column; returns set_separate of the sorted
// keys when `separate_out` is set
void mg_sort(int ktype, const void* col, size_t* idx, size_t n, std::vector<size_t>* sep) {
  dispatch_key(ktype, [&](auto z) {
    using K = decltype(z);
    const K* c = reinterpret_cast<const K*>(col);
    std::vector<K> v(n);
    for (size_t i = 0; i < n; i++) v[i] = c[idx[i]];
    sort_kv(v.data(), idx, n, false);
    if (sep) {
      sep->resize(n + 1);
      sep->resize(separate(v.data(), n, sep->data()));
    }
  });
}

// set_separate of col[idx] without sorting (multi_group_by_split)
std::vector<size_t> mg_split(int ktype, const void* col, const size_t* idx, size_t n) {
  std::vector<size_t> sep(n + 1);
  dispatch_key(ktype, [&](auto z) {
    using K = decltype(z);
    const K* c = reinterpret_cast<const K*>(col);
    std::vector<K> v(n);
    for (size_t i = 0; i < n; i++) v[i] = c[idx[i]];
    sep.resize(separate(v.data(), n, sep.data()));
  });
  return sep;
}

size_t multi_group_by_t(int ncols, const int* ktypes, const void* const* cols, size_t* idx,
                        size_t n, size_t* split, void* const* keys_out) {
  for (int i = 0; i < ncols - 1; i++)
    mg_sort(ktypes[ncols - i - 1], cols[ncols - i - 1], idx, n, nullptr);
  std::vector<size_t> sp;
  mg_sort(ktypes[0], cols[0], idx, n, &sp);
  for (int i = 1; i < ncols; i++) {
    std::vector<size_t> t = mg_split(ktypes[i], cols[i], idx, n);
    std::vector<size_t> u;
    std::set_union(sp.begin(), sp.end(), t.begin(), t.end(), std::back_inserter(u));
    sp.swap(u);
  }
  for (int c = 0; c < ncols; c++) {
    dispatch_key(ktypes[c], [&](auto z) {
      using K = decltype(z);
      const K* col = reinterpret_cast<const K*>(cols[c]);
      K* ko = reinterpret_cast<K*>(keys_out[c]);
      for (size_t g = 0; g + 1 < sp.size(); g++) ko[g] = col[idx[sp[g]]];
    });
  }
  std::copy(sp.begin(), sp.end(), split);
  return sp.size();
}

thread_local std::string g_err;

}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

int orc_radix_sort(int ktype, void* keys, int vbytes, void* vals, size_t n,
                   int positive_only, int descending) {
  uint8_t* v = reinterpret_cast<uint8_t*>(vals);
  if (vbytes != 0 && vbytes != 4 && vbytes != 8) return -1;
  bool po = positive_only != 0, de = descending != 0;
  switch (ktype) {
    case T_I32: radix_sort_int<int32_t, uint32_t>((int32_t*)keys, v, vbytes, n, po, de); break;
    case T_U32: radix_sort_int<uint32_t, uint32_t>((uint32_t*)keys, v, vbytes, n, po, de); break;
    case T_I64: radix_sort_int<int64_t, uint64_t>((int64_t*)keys, v, vbytes, n, po, de); break;
    case T_U64: radix_sort_int<uint64_t, uint64_t>((uint64_t*)keys, v, vbytes, n, po, de); break;
    case T_F32: radix_sort_fp<float, uint32_t>((float*)keys, v, vbytes, n, po, de); break;
    case T_F64: radix_sort_fp<double, uint64_t>((double*)keys, v, vbytes, n, po, de); break;
    default: return -1;
  }
  return 0;
}

// prefix_sum x86 body, core/prefix_sum.hpp:67-76: sequential inclusive scan.
int orc_prefix_sum(int type, const void* in, void* out, size_t n) {
  try {
    dispatch_key(type, [&](auto z) {
      using T = decltype(z);
      const T* a = reinterpret_cast<const T*>(in);
      T* o = reinterpret_cast<T*>(out);
      T cur = 0;
      for (size_t j = 0; j < n; j++) { cur = a[j] + cur; o[j] = cur; }
    });
  } catch (std::exception& e) { g_err = e.what(); return -1; }
  return 0;
}

// returns the number of entries written to out (capacity n + 1)
size_t orc_set_separate(int type, const void* keys, size_t n, size_t* out) {
  size_t r = 0;
  dispatch_key(type, [&](auto z) {
    using T = decltype(z);
    r = separate(reinterpret_cast<const T*>(keys), n, out);
  });
  return r;
}

int orc_set_is_unique(int type, const void* keys, size_t n) {
  int r = 1;
  dispatch_key(type, [&](auto z) {
    using T = decltype(z);
    r = is_unique(reinterpret_cast<const T*>(keys), n);
  });
  return r;
}

void orc_lower_bound(int type, const void* sorted, size_t m, const void* q,
                     size_t nq, size_t* out, int upper) {
  dispatch_key(type, [&](auto z) {
    using T = decltype(z);
    if (upper) upper_bound_t(reinterpret_cast<const T*>(sorted), m,
                             reinterpret_cast<const T*>(q), nq, out);
    else lower_bound_t(reinterpret_cast<const T*>(sorted), m,
                       reinterpret_cast<const T*>(q), nq, out);
  });
}

void* orc_ht_build(int ktype, const void* keys, const uint64_t* vals, size_t n,
                   int check_unique, int* is_unique_ok) {
  HT* h = new HT;
  h->ktype = ktype;
  h->kbytes = (ktype == T_I32 || ktype == T_U32 || ktype == T_F32) ? 4 : 8;
  try {
    dispatch_key(ktype, [&](auto z) {
      using K = decltype(z);
      ht_build_t<K>(h, keys, vals, n, check_unique, is_unique_ok);
    });
  } catch (std::exception& e) { g_err = e.what(); delete h; return nullptr; }
  return h;
}

size_t orc_ht_table_size(void* hp) { return reinterpret_cast<HT*>(hp)->table_size; }
size_t orc_ht_conflict_size(void* hp) { return reinterpret_cast<HT*>(hp)->ckey.size(); }

size_t orc_ht_lookup(void* hp, const void* keys, size_t n, uint64_t* out,
                     size_t* miss) {
  HT* h = reinterpret_cast<HT*>(hp);
  size_t r = 0;
  dispatch_key(h->ktype, [&](auto z) {
    using K = decltype(z);
    r = ht_lookup_t<K>(h, keys, n, out, miss);
  });
  return r;
}

void orc_ht_check_existence(void* hp, const void* keys, size_t n, int* out) {
  HT* h = reinterpret_cast<HT*>(hp);
  dispatch_key(h->ktype, [&](auto z) {
    using K = decltype(z);
    ht_exist_t<K>(h, keys, n, out);
  });
}

// all_keys (hashtable.hpp:1584-1607); out capacity = number of build rows
size_t orc_ht_all_keys(void* hp, void* out) {
  HT* h = reinterpret_cast<HT*>(hp);
  size_t c = 0;
  for (size_t s = 0; s < h->table_size; s++) {
    if (!h->filled[s]) continue;
    if (h->kbytes == 4) reinterpret_cast<uint32_t*>(out)[c++] = uint32_t(h->key[s]);
    else reinterpret_cast<uint64_t*>(out)[c++] = h->key[s];
  }
  for (size_t i = 0; i < h->ckey.size(); i++) {
    if (h->kbytes == 4) reinterpret_cast<uint32_t*>(out)[c++] = uint32_t(h->ckey[i]);
    else reinterpret_cast<uint64_t*>(out)[c++] = h->ckey[i];
  }
  return c;
}

void orc_ht_free(void* hp) { delete reinterpret_cast<HT*>(hp); }

// Returns the pair count; *lo/*ro are malloc'ed (free with orc_free).
size_t orc_equi_join(int ktype, const void* left, const size_t* lidx, size_t nl,
                     const void* right, const size_t* ridx, size_t nr, int algo,
                     size_t** lo, size_t** ro, int* used_hash) {
  std::vector<size_t> a, b;
  dispatch_key(ktype, [&](auto z) {
    using K = decltype(z);
    equi_join_t<K>(left, lidx, nl, right, ridx, nr, algo, a, b, used_hash, nullptr);
  });
  size_t m = a.size();
  *lo = static_cast<size_t*>(std::malloc((m ? m : 1) * sizeof(size_t)));
  *ro = static_cast<size_t*>(std::malloc((m ? m : 1) * sizeof(size_t)));
  if (m) {
    std::memcpy(*lo, a.data(), m * sizeof(size_t));
    std::memcpy(*ro, b.data(), m * sizeof(size_t));
  }
  return m;
}

void orc_free(void* p) { std::free(p); }

// split capacity n + 1, keys_out capacity n; returns split length.
size_t orc_group_by(int ktype, const void* col, size_t* idx, size_t n,
                    size_t* split, void* keys_out) {
  size_t r = 0;
  dispatch_key(ktype, [&](auto z) {
    using K = decltype(z);
    r = group_by_t<K>(col, idx, n, split, keys_out);
  });
  return r;
}

int orc_grouped_agg(int op, int vtype, const void* col, const size_t* gidx,
                    size_t n, const size_t* split, size_t ns,
                    const size_t* nulls, size_t nn, void* out) {
  try {
    dispatch_key(vtype, [&](auto z) {
      using T = decltype(z);
      grouped_agg_t<T>(op, col, gidx, n, split, ns, nulls, nn, out);
    });
  } catch (std::exception& e) { g_err = e.what(); return -1; }
  return 0;
}


// Outer join: returns the pair count; *lo/*ro/*only are malloc'ed (free with
// orc_free).  algo 0 VE, 1 x86, 2 VE sort-merge branch.
size_t orc_outer_equi_join(int ktype, const void* left, const size_t* lidx, size_t nl,
                           const void* right, const size_t* ridx, size_t nr, int algo,
                           size_t** lo, size_t** ro, size_t** only, size_t* n_only,
                           int* used_hash) {
  std::vector<size_t> a, b, c;
  dispatch_key(ktype, [&](auto z) {
    using K = decltype(z);
    equi_join_t<K>(left, lidx, nl, right, ridx, nr, algo, a, b, used_hash, &c);
  });
  auto out = [](const std::vector<size_t>& v) {
    size_t* p = static_cast<size_t*>(std::malloc((v.empty() ? 1 : v.size()) * sizeof(size_t)));
    if (!v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(size_t));
    return p;
  };
  *lo = out(a);
  *ro = out(b);
  *only = out(c);
  *n_only = c.size();
  return a.size();
}

int orc_gather(int type, const void* src, size_t src_n, const size_t* idx, size_t n, void* dst) {
  int rc = 0;
  try {
    dispatch_key(type, [&](auto z) {
      using T = decltype(z);
      rc = gather_t(reinterpret_cast<const T*>(src), src_n, idx, n, reinterpret_cast<T*>(dst));
    });
  } catch (std::exception& e) { g_err = e.what(); return -1; }
  if (rc) g_err = "gather: row index out of range";
  return rc;
}

size_t orc_gather_nulls(const size_t* idx, size_t n, const size_t* nulls, size_t nn,
                        size_t* out) {
  return gather_nulls_t(idx, n, nulls, nn, out);
}

// out capacity n; returns the count, or (size_t)-1 on a bad code
size_t orc_filter(int ctype, const void* col, size_t n, int op, int stype, const void* scalar,
                  const size_t* nulls, size_t nn, size_t* out) {
  size_t r = 0;
  try {
    dispatch_key(ctype, [&](auto zc) {
      using T = decltype(zc);
      dispatch_key(stype, [&](auto zs) {
        using U = decltype(zs);
        U c;
        std::memcpy(&c, scalar, sizeof(U));
        r = filter_t(reinterpret_cast<const T*>(col), n, op, c, nulls, nn, out);
      });
    });
  } catch (std::exception& e) { g_err = e.what(); return ~size_t(0); }
  return r;
}

size_t orc_filter_pair(int atype, const void* a, int btype, const void* b, size_t n, int op,
                       const size_t* an, size_t na, const size_t* bn, size_t nb, size_t* out) {
  size_t r = 0;
  try {
    dispatch_key(atype, [&](auto za) {
      using T = decltype(za);
      dispatch_key(btype, [&](auto zb) {
        using U = decltype(zb);
        r = filter_pair_t(reinterpret_cast<const T*>(a), reinterpret_cast<const U*>(b), n, op,
                          an, na, bn, nb, out);
      });
    });
  } catch (std::exception& e) { g_err = e.what(); return ~size_t(0); }
  return r;
}

int orc_sort_index(int ktype, const void* col, const size_t* idx_in, size_t n, int desc,
                   void* sorted, size_t* idx_out) {
  try {
    dispatch_key(ktype, [&](auto z) {
      using K = decltype(z);
      sort_index_t(reinterpret_cast<const K*>(col), idx_in, n, desc != 0,
                   reinterpret_cast<K*>(sorted), idx_out);
    });
  } catch (std::exception& e) { g_err = e.what(); return -1; }
  return 0;
}

// split capacity n + 1, keys_out[c] capacity n; returns the split length
size_t orc_multi_group_by(int ncols, const int* ktypes, const void* const* cols, size_t* idx,
                          size_t n, size_t* split, void* const* keys_out) {
  try {
    return multi_group_by_t(ncols, ktypes, cols, idx, n, split, keys_out);
  } catch (std::exception& e) { g_err = e.what(); return 0; }
}

// star join: returns the row count (or (size_t)-1 with the message in
// orc_last_error); left_out and right_out[d] hold nl entries.
size_t orc_star_join(int nd, const int* ktypes, const void* const* left_cols,
                     const size_t* lidx, size_t nl, const void* const* right_cols,
                     const size_t* const* ridx, const size_t* nr, size_t* left_out,
                     size_t* const* right_out) {
  try {
    std::vector<size_t> lrows(nl);
    for (size_t i = 0; i < nl; i++) lrows[i] = lidx ? lidx[i] : i;
    std::vector<std::vector<size_t>> rids;
    for (int d = 0; d < nd; d++) {
      dispatch_key(ktypes[d], [&](auto z) {
        using K = decltype(z);
        star_dim_t<K>(left_cols[d], right_cols[d], ridx[d], nr[d], lrows, rids);
      });
    }
    std::copy(lrows.begin(), lrows.end(), left_out);
    for (int d = 0; d < nd; d++) std::copy(rids[d].begin(), rids[d].end(), right_out[d]);
    return lrows.size();
  } catch (std::exception& e) { g_err = e.what(); return ~size_t(0); }
}

}  // extern "C"
