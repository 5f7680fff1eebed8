// Internal interface of the gfx950 unique_hashtable (shared with the join).
#pragma once

#include "fhip_common.hpp"

// One slot of the main table: the owner's key bits (4-byte keys in the low
// half), its value and the owner row + 1 (0 = empty) in one 32-byte sector,
// so that a probe reads one sector.
struct alignas(32) fhip_ht_slot {
  unsigned long long key;
  unsigned long long val;
  unsigned owner;
  unsigned pad0;
  unsigned long long pad1;
};

// Secondary index over the sorted conflict list (integer keys): open
// addressing by a 64-bit mix of the key; every distinct conflict key once,
// with the value of its first row in the sorted list -- what lower_bound
// into the list returns (dataframe/hashtable.hpp:1442-1515).
struct alignas(32) fhip_ht_cslot {
  unsigned long long key;
  unsigned long long val;
  unsigned long long tag;  // conflict index + 1, 0 = empty
  unsigned long long pad;
};

struct fhip_ht {
  int ktype = 0;
  int has_vals = 0;
  size_t n = 0;           // build rows
  size_t table_size = 0;  // prime >= 3n (dataframe/hashtable.hpp:115-152)
  fhip_ht_slot* slots = nullptr;  // [table_size]
  void* conf_key = nullptr;       // [nconf] conflict keys, sorted (stable)
  uint64_t* conf_val = nullptr;   // [nconf]
  size_t nconf = 0;
  fhip_ht_cslot* cidx = nullptr;  // [cmask + 1] secondary index (integer keys), or null
  size_t cmask = 0;
};

namespace fhip {

// Builds the table with the reference's layout; *is_unique_ok follows the
// (k, v, int&) constructor when check_unique is set.
fhip_ht* ht_build(int ktype, const void* keys, const uint64_t* vals, size_t n, bool check_unique,
                  int* is_unique_ok, hipStream_t st);
void ht_destroy(fhip_ht* h);

// Per query: found flag (1/0) and the value (0 when missing).
void ht_probe(const fhip_ht* h, const void* q, size_t m, uint64_t* vals_out, uint8_t* found,
              hipStream_t st);

size_t table_size_for(size_t n);

}  // namespace fhip
