// Internal interface of the gfx950 radix sort, shared by group_by, the
// joins, the hash table conflict list and the batched/learned sorts.
#pragma once

#include "fhip_common.hpp"

namespace fhip {

// How the key bits are mapped to an unsigned order-preserving value before
// digits are extracted (the reference's flips, core/radix_sort.hpp:155-195).
enum KeyMode : int {
  KM_RAW = 0,        // unsigned keys, or positive_only, or floats w/o negatives
  KM_FLIPTOP = 1,    // signed ints: xor the sign bit
  KM_FLOATFLIP = 2,  // IEEE keys with a negative: total-order flip
};

struct SortPlan {
  int kbytes = 8;
  int mode = KM_RAW;
  int desc = 0;
  std::vector<int> passes;  // byte positions to run, ascending
};

// Stable sort of n keys (+ optional payload of vbytes bytes) from
// (kin, vin) into (kout, vout).  kout may equal kin (in place).  Follows the
// reference semantics of radix_sort(K*, V*, size_t, bool) including the
// positive_only rule and the < 64 insertion path.  low_bytes >= 0 (integer
// keys, positive_only false): the caller knows that every key has the same
// bytes above the lowest low_bytes (in the sort's order, the sign bit of
// signed keys flipped), so those passes run without the host reading back
// the histograms.
// A caller-provided device region that temporaries are carved from instead
// of the workspace pool (no allocation); take() throws when it runs out.
struct Arena {
  char* base = nullptr;
  size_t cap = 0, used = 0;
  void* take(size_t bytes) {
    const size_t at = (used + 255) & ~size_t(255);
    FHIP_REQUIRE(at + bytes <= cap, FHIP_ERR_TOO_LARGE, "radix_sort: scratch region too small");
    used = at + bytes;
    return base + at;
  }
};

void sort_pairs(int ktype, const void* kin, void* kout, int vbytes,
                const void* vin, void* vout, size_t n, bool positive_only,
                bool descending, hipStream_t stream, int low_bytes = -1,
                Arena* arena = nullptr);

// Bytes of scratch sort_pairs needs with an Arena (in place, one temporary
// key/payload buffer, the histograms and the look-back words).
size_t sort_scratch_bytes(size_t n, int kbytes, int vbytes);

// Plain stable ascending sort of unsigned 64-bit (or 32-bit) keys that only
// uses the bytes in [0, nbytes); for internal callers that already know the
// key range.  No insertion path, no flips.
void sort_pairs_bits(int kbytes, const void* kin, void* kout, int vbytes,
                     const void* vin, void* vout, size_t n, int key_bits,
                     hipStream_t stream);

// Narrow stable sort (the sort-merge join's): integer keys whose bits above
// the low 32 -- in the sort's order, the sign bit of signed keys flipped --
// are shared by every key, with row ids (ids, size_t) that are consecutive.
// Writes each key's low 32 bits in the sort's order (k32, ascending as
// unsigned) and its row 0..n-1 (row): half the bytes of a key/row-id sort
// per pass.  Returns false, having written nothing, when the keys differ
// above their low 32 bits; *hi gets the shared high bits (in the sort's
// order).  The row ids are checked on the device: *iota_bad (a device word
// the caller zeroes) becomes non-zero when ids[i] != ids[0] + i for some i,
// and the result must then not be used.
bool sort_narrow(int ktype, const void* keys, const size_t* ids, size_t n, uint32_t* k32,
                 uint32_t* row, unsigned* hi, unsigned* iota_bad, hipStream_t st);

// Graph-capturable variant (no host synchronisation, memset or allocation):
// stable sort of unsigned kbytes-wide keys on bytes [0, nbytes) into
// (kout, vout), using a caller-provided workspace of
// sort_async_workspace_bytes(n, kbytes, vbytes) bytes.
size_t sort_async_workspace_bytes(size_t n, int kbytes, int vbytes);
void sort_pairs_async(int kbytes, int nbytes, const void* kin, void* kout, int vbytes,
                      const void* vin, void* vout, size_t n, void* ws, hipStream_t st);

}  // namespace fhip
