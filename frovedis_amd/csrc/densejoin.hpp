// The dense-key hash join (densejoin.hip), called by equi_join for integer
// keys before its open-addressing table.
#pragma once

#include "fhip_common.hpp"

namespace fhip {

constexpr int DJ_NA = 0;    // the dense table does not apply: use another table
constexpr int DJ_DUP = 1;   // RIGHT repeats a key (not unique)
constexpr int DJ_DONE = 2;  // joined

// Inner join: the pairs (lidx of a left row, right row id of its partner)
// in left order, at most `capacity` written, *npairs = all of them.
int dense_join_pairs(int ktype, const void* left, const size_t* lidx, size_t nl,
                     const void* right, const size_t* ridx, size_t nr, size_t* lo, size_t* ro,
                     size_t capacity, size_t* npairs, hipStream_t st);

// Per left row: found (nl bytes, 0/1) and the partner's right row id (vals,
// nl 64-bit words, 0 when none), for the outer join; both are allocated
// here when the dense table applies.
int dense_join_flags(int ktype, const void* left, size_t nl, const void* right,
                     const size_t* ridx, size_t nr, DBuf& vals, DBuf& found, hipStream_t st);

}  // namespace fhip
