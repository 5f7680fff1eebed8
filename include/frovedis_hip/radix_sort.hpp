// Drop-in for frovedis core/radix_sort.hpp: the same free templates and
// signatures, forwarding the supported key/payload types to
// libfrovedis_hip (fhip_radix_sort, host pointers, sorted in place).
//   radix_sort(K*, V*, size_t, bool positive_only = false)    :197-219
//   radix_sort(K*, size_t, bool positive_only = false)        :221-242
//   radix_sort_desc(K*, V*, size_t, bool) / (K*, size_t, bool) :342-487
//   radix_sort(std::vector<K>&, std::vector<V>&, bool)        :489-496
//   radix_sort(std::vector<K>&, bool)                          :498-502
//   radix_sort_desc(std::vector<...>&...)                      :504-518
// K: int, unsigned int, long, unsigned long, long long, unsigned long long,
// float, double.  V: any 4- or 8-byte trivially copyable type.
#ifndef FROVEDIS_HIP_RADIX_SORT_HPP
#define FROVEDIS_HIP_RADIX_SORT_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../frovedis_hip.h"

namespace frovedis {
namespace fhip_detail {

template <class K>
constexpr int key_code() {
  static_assert(std::is_integral<K>::value || std::is_floating_point<K>::value,
                "radix_sort: key must be an integer or floating point type");
  static_assert(sizeof(K) == 4 || sizeof(K) == 8, "radix_sort: key must be 4 or 8 bytes");
  return std::is_floating_point<K>::value ? (sizeof(K) == 4 ? FHIP_F32 : FHIP_F64)
         : std::is_signed<K>::value        ? (sizeof(K) == 4 ? FHIP_I32 : FHIP_I64)
                                           : (sizeof(K) == 4 ? FHIP_U32 : FHIP_U64);
}

inline void check(int rc) {
  if (rc != FHIP_OK) throw std::runtime_error(fhip_last_error());
}

template <class K, class V>
void sort_kv(K* k, V* v, size_t size, bool positive_only, bool desc) {
  static_assert(sizeof(V) == 4 || sizeof(V) == 8, "radix_sort: value must be 4 or 8 bytes");
  static_assert(std::is_trivially_copyable<V>::value, "radix_sort: value must be trivially copyable");
  check(fhip_radix_sort(key_code<K>(), k, int(sizeof(V)), v, size, positive_only ? 1 : 0,
                        desc ? 1 : 0));
}

template <class K>
void sort_k(K* k, size_t size, bool positive_only, bool desc) {
  check(fhip_radix_sort(key_code<K>(), k, 0, nullptr, size, positive_only ? 1 : 0, desc ? 1 : 0));
}

}  // namespace fhip_detail

template <class K, class V>
void radix_sort(K* key_array, V* val_array, size_t size, bool positive_only = false) {
  fhip_detail::sort_kv(key_array, val_array, size, positive_only, false);
}

template <class K>
void radix_sort(K* key_array, size_t size, bool positive_only = false) {
  fhip_detail::sort_k(key_array, size, positive_only, false);
}

template <class K, class V>
void radix_sort_desc(K* key_array, V* val_array, size_t size, bool positive_only = false) {
  fhip_detail::sort_kv(key_array, val_array, size, positive_only, true);
}

template <class K>
void radix_sort_desc(K* key_array, size_t size, bool positive_only = false) {
  fhip_detail::sort_k(key_array, size, positive_only, true);
}

template <class K, class V>
void radix_sort(std::vector<K>& key_array, std::vector<V>& val_array,
                bool positive_only = false) {
  size_t size = key_array.size();
  if (val_array.size() != size)
    throw std::runtime_error("radix_sort: different size of arrays");
  radix_sort(key_array.data(), val_array.data(), size, positive_only);
}

template <class K>
void radix_sort(std::vector<K>& key_array, bool positive_only = false) {
  radix_sort(key_array.data(), key_array.size(), positive_only);
}

template <class K, class V>
void radix_sort_desc(std::vector<K>& key_array, std::vector<V>& val_array,
                     bool positive_only = false) {
  size_t size = key_array.size();
  if (val_array.size() != size)
    throw std::runtime_error("radix_sort: different size of arrays");
  radix_sort_desc(key_array.data(), val_array.data(), size, positive_only);
}

template <class K>
void radix_sort_desc(std::vector<K>& key_array, bool positive_only = false) {
  radix_sort_desc(key_array.data(), key_array.size(), positive_only);
}

}  // namespace frovedis

#endif
