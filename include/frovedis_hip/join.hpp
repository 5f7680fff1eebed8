// Drop-in for frovedis dataframe/join.hpp equi_join<T> (:304-326) and the
// per-rank group-by / grouped aggregation lambdas of
// dataframe/dfcolumn_groupby.cc:25-58 and dfcolumn_grouped_agg.cc
// (sum_helper :105-190, max_helper :366-452, min_helper, count_helper
// :747-759), with one node (hash_divide and merge_map are identities unless
// the caller passes its own hash_divide).
#ifndef FROVEDIS_HIP_JOIN_HPP
#define FROVEDIS_HIP_JOIN_HPP

#include <vector>

#include "hashtable.hpp"
#include "set_operations.hpp"

namespace frovedis {

// Pairs in the reference's VE orders: probe (left) order when the right
// keys are unique, (key, left, right) order otherwise.  Like the VE body,
// the inputs are left as given except that the sort-merge path sorts left /
// right (with their indices) in place.
template <class T>
void equi_join(std::vector<T>& left, std::vector<size_t>& left_idx, std::vector<T>& right,
               std::vector<size_t>& right_idx, std::vector<size_t>& left_idx_out,
               std::vector<size_t>& right_idx_out) {
  if (left.size() != left_idx.size() || right.size() != right_idx.size())
    throw std::runtime_error("sizes of key and value are not the same");
  size_t cap = left.size() ? left.size() : 1;
  for (int attempt = 0; attempt < 2; attempt++) {
    left_idx_out.resize(cap);
    right_idx_out.resize(cap);
    size_t nout = 0;
    int used_hash = 0;
    // the sort-merge path sorts the inputs in place, as the VE body does;
    // they come back with the pairs (nothing comes back when the pairs did
    // not fit, so the retry starts from the inputs as given)
    int rc = fhip_equi_join_inplace(fhip_detail::key_code<T>(), left.data(), left_idx.data(),
                                    left.size(), right.data(), right_idx.data(), right.size(),
                                    FHIP_JOIN_AUTO, left_idx_out.data(), right_idx_out.data(),
                                    cap, &nout, &used_hash);
    if (rc == FHIP_ERR_TOO_LARGE && nout > cap) {
      cap = nout;
      continue;
    }
    fhip_detail::check(rc);
    left_idx_out.resize(nout);
    right_idx_out.resize(nout);
    return;
  }
  throw std::runtime_error("equi_join: output did not fit");
}

// First lambda of group_by_impl on one rank: newval = val[idx];
// radix_sort(newval, idx); split = set_separate(newval); group keys.
template <class T>
std::vector<T> local_group_by(const std::vector<T>& val, std::vector<size_t>& idx,
                              std::vector<size_t>& split) {
  split.resize(idx.size() + 1);
  std::vector<T> keys(idx.size() ? idx.size() : 1);
  size_t ns = 0;
  fhip_detail::check(fhip_group_by(fhip_detail::key_code<T>(), val.data(), val.size(), idx.data(),
                                   idx.size(), split.data(), &ns, keys.data()));
  split.resize(ns);
  keys.resize(ns - 1);
  return keys;
}

namespace fhip_detail {
template <class R, class T>
std::vector<std::vector<R>> grouped(int op, std::vector<T>& org_val,
                                    std::vector<size_t>& grouped_idx,
                                    std::vector<size_t>& idx_split,
                                    std::vector<std::vector<size_t>>& hash_divide,
                                    std::vector<size_t>& nulls) {
  const size_t G = idx_split.size() ? idx_split.size() - 1 : 0;
  std::vector<R> ret(G);
  check(fhip_grouped_agg(op, key_code<T>(), org_val.data(), org_val.size(), grouped_idx.data(),
                         grouped_idx.size(), idx_split.data(), idx_split.size(), nulls.data(),
                         nulls.size(), ret.data()));
  std::vector<std::vector<R>> hashed(hash_divide.size());
  for (size_t i = 0; i < hash_divide.size(); i++) {
    hashed[i].resize(hash_divide[i].size());
    for (size_t j = 0; j < hash_divide[i].size(); j++) hashed[i][j] = ret[hash_divide[i][j]];
  }
  return hashed;
}
}  // namespace fhip_detail

template <class T>
std::vector<std::vector<T>> sum_helper(std::vector<T>& org_val, std::vector<size_t>& grouped_idx,
                                       std::vector<size_t>& idx_split,
                                       std::vector<std::vector<size_t>>& hash_divide,
                                       std::vector<size_t>& nulls) {
  return fhip_detail::grouped<T>(FHIP_AGG_SUM, org_val, grouped_idx, idx_split, hash_divide, nulls);
}

template <class T>
std::vector<std::vector<T>> max_helper(std::vector<T>& org_val, std::vector<size_t>& grouped_idx,
                                       std::vector<size_t>& idx_split,
                                       std::vector<std::vector<size_t>>& hash_divide,
                                       std::vector<size_t>& nulls) {
  return fhip_detail::grouped<T>(FHIP_AGG_MAX, org_val, grouped_idx, idx_split, hash_divide, nulls);
}

template <class T>
std::vector<std::vector<T>> min_helper(std::vector<T>& org_val, std::vector<size_t>& grouped_idx,
                                       std::vector<size_t>& idx_split,
                                       std::vector<std::vector<size_t>>& hash_divide,
                                       std::vector<size_t>& nulls) {
  return fhip_detail::grouped<T>(FHIP_AGG_MIN, org_val, grouped_idx, idx_split, hash_divide, nulls);
}

template <class T>
std::vector<std::vector<size_t>> count_helper(std::vector<T>& val,
                                              std::vector<size_t>& grouped_idx,
                                              std::vector<size_t>& idx_split,
                                              std::vector<std::vector<size_t>>& hash_divide,
                                              std::vector<size_t>& nulls) {
  return fhip_detail::grouped<size_t>(FHIP_AGG_COUNT, val, grouped_idx, idx_split, hash_divide,
                                      nulls);
}

}  // namespace frovedis

#endif
