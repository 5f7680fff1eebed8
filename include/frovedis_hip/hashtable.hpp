// Drop-in for frovedis dataframe/hashtable.hpp: class template
// unique_hashtable<K, V> with the reference's public members (:57-110),
// backed by a table in HBM (fhip_ht_* entry points).  V must be 8 bytes
// (the reference instantiates it with size_t).  Copies share the table.
#ifndef FROVEDIS_HIP_HASHTABLE_HPP
#define FROVEDIS_HIP_HASHTABLE_HPP

#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

#include "radix_sort.hpp"

namespace frovedis {

template <class K>
inline size_t myhash(const K& key, size_t size) {  // dataframe/hashtable.hpp:28-54
  unsigned long ukey = 0;
  static_assert(sizeof(K) == 4 || sizeof(K) == 8, "myhash: 4 or 8 byte keys");
  if (sizeof(K) == 4) {
    unsigned int u;
    std::memcpy(&u, &key, 4);
    ukey = u;
  } else {
    std::memcpy(&ukey, &key, 8);
  }
  return ukey % size;
}

template <class K, class V>
class unique_hashtable {
  static_assert(sizeof(V) == 8, "unique_hashtable: value must be 8 bytes (size_t)");

 public:
  unique_hashtable() {}
  unique_hashtable(const std::vector<K>& k, const std::vector<V>& v) { build(k, &v, false, nullptr); }
  // for creating set (used hash join), accept duplicated keys
  unique_hashtable(const std::vector<K>& k) { build(k, nullptr, false, nullptr); }
  unique_hashtable(const std::vector<K>& k, const std::vector<V>& v, int& is_unique_ok) {
    build(k, &v, true, &is_unique_ok);
  }

  std::vector<K> all_keys() {
    std::vector<K> ret(n_);
    size_t m = 0;
    if (ht_) fhip_detail::check(fhip_ht_all_keys(ht_.get(), ret.data(), &m));
    ret.resize(m);
    return ret;
  }
  std::vector<V> lookup(const std::vector<K>& k) {
    std::vector<V> v(k.size());
    if (ht_)
      fhip_detail::check(fhip_ht_lookup(ht_.get(), k.data(), k.size(),
                                        reinterpret_cast<uint64_t*>(v.data()), nullptr, nullptr));
    return v;
  }
  std::vector<V> lookup(const std::vector<K>& k, std::vector<size_t>& miss) {
    std::vector<V> v(k.size());
    miss.resize(k.size());
    size_t nm = k.size();
    if (ht_)
      fhip_detail::check(fhip_ht_lookup(ht_.get(), k.data(), k.size(),
                                        reinterpret_cast<uint64_t*>(v.data()), miss.data(), &nm));
    else
      for (size_t i = 0; i < k.size(); i++) miss[i] = i;
    miss.resize(nm);
    return v;
  }
  std::vector<int> check_existence(const std::vector<K>& k) {
    std::vector<int> r(k.size(), 0);
    if (ht_) fhip_detail::check(fhip_ht_check_existence(ht_.get(), k.data(), k.size(), r.data()));
    return r;
  }
  void clear() { ht_.reset(); }  // to save memory

 private:
  void build(const std::vector<K>& k, const std::vector<V>* v, bool check_unique, int* uniq) {
    if (v && v->size() != k.size())
      throw std::runtime_error("sizes of key and value are different");
    fhip_ht* h = nullptr;
    int flag = 1;
    fhip_detail::check(fhip_ht_build(fhip_detail::key_code<K>(), k.data(),
                                     v ? reinterpret_cast<const uint64_t*>(v->data()) : nullptr,
                                     k.size(), check_unique ? 1 : 0, &flag, &h));
    ht_.reset(h, [](fhip_ht* p) { fhip_ht_destroy(p); });
    n_ = k.size();
    if (uniq) *uniq = flag;
  }
  std::shared_ptr<fhip_ht> ht_;
  size_t n_ = 0;
};

}  // namespace frovedis

#endif
