// Internal interface of the gfx950 unique_hashtable (shared with the join).
#pragma once

#include "fhip_common.hpp"

struct fhip_ht {
  int ktype = 0;
  int has_vals = 0;
  size_t n = 0;           // build rows
  size_t table_size = 0;  // prime >= 3n (dataframe/hashtable.hpp:115-152)
  unsigned* owner = nullptr;     // [table_size] row index + 1 of the slot owner, 0 = empty
  void* slot_key = nullptr;      // [table_size] key of the owner (valid where owner != 0)
  uint64_t* slot_val = nullptr;  // [table_size]
  void* conf_key = nullptr;      // [nconf] conflict keys, sorted (stable)
  uint64_t* conf_val = nullptr;  // [nconf]
  size_t nconf = 0;
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
