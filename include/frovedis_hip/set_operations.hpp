// Drop-in for the parts of frovedis core/prefix_sum.hpp and
// core/set_operations.hpp on the dataframe path:
//   prefix_sum(const T*, T*, size_t), prefix_sum(const std::vector<T>&)
//     (prefix_sum.hpp:67-88)
//   set_separate(const T*, size_t), set_separate(const std::vector<T>&)
//     (set_operations.hpp:146-163, :1769-1772)
//   set_unique / set_is_unique (set_operations.hpp:166-197)
// forwarding to libfrovedis_hip with host pointers.
#ifndef FROVEDIS_HIP_SET_OPERATIONS_HPP
#define FROVEDIS_HIP_SET_OPERATIONS_HPP

#include <cstddef>
#include <vector>

#include "radix_sort.hpp"

namespace frovedis {

template <class T>
void prefix_sum(const T* valp, T* outp, size_t size) {
  fhip_detail::check(fhip_prefix_sum(fhip_detail::key_code<T>(), valp, outp, size));
}

template <class T>
std::vector<T> prefix_sum(const std::vector<T>& val) {
  size_t size = val.size();
  if (size == 0) return std::vector<T>();
  std::vector<T> out(size);
  prefix_sum(val.data(), out.data(), size);
  return out;
}

template <class T>
std::vector<size_t> set_separate(const T* keyp, size_t size) {
  std::vector<size_t> ret(size + 1);
  size_t got = 0;
  fhip_detail::check(fhip_set_separate(fhip_detail::key_code<T>(), keyp, size, ret.data(), &got));
  ret.resize(got);
  return ret;
}

template <class T>
std::vector<size_t> set_separate(const std::vector<T>& key) {
  return set_separate(key.data(), key.size());
}

template <class T>
std::vector<T> set_unique(const T* keyp, size_t size) {
  auto sep = set_separate(keyp, size);
  std::vector<T> ret(sep.size() ? sep.size() - 1 : 0);
  for (size_t i = 0; i + 1 < sep.size(); i++) ret[i] = keyp[sep[i]];
  return ret;
}

template <class T>
std::vector<T> set_unique(const std::vector<T>& key) {
  return set_unique(key.data(), key.size());
}

template <class T>
int set_is_unique(const T* keyp, size_t size) {
  int u = 1;
  fhip_detail::check(fhip_set_is_unique(fhip_detail::key_code<T>(), keyp, size, &u));
  return u;
}

template <class T>
int set_is_unique(const std::vector<T>& key) {
  return set_is_unique(key.data(), key.size());
}

}  // namespace frovedis

#endif
