// The dense-key hash join (densejoin.hip), called by equi_join for integer
// keys before its open-addressing table.
#pragma once

#include "fhip_common.hpp"

namespace fhip {

constexpr int DJ_NA = 0;    // the dense table does not apply: use another table
constexpr int DJ_DUP = 1;   // RIGHT repeats a key (not unique)
constexpr int DJ_DONE = 2;  // joined

// Inner join: the pairs (lidx of a left row, right row id of its partner)
// in left order, at most `capacity` written, *npairs = all of them.  With
// only_left (nl entries; the LEFT outer join) the left row ids without a
// partner too, in left order, *n_only of them.
int dense_join_pairs(int ktype, const void* left, const size_t* lidx, size_t nl,
                     const void* right, const size_t* ridx, size_t nr, size_t* lo, size_t* ro,
                     size_t capacity, size_t* npairs, size_t* only_left, size_t* n_only,
                     hipStream_t st);

}  // namespace fhip
