// Exercises the drop-in headers the way the reference's callers use them
// (std::vector storage, frovedis:: names) and checks the results against
// std::stable_sort, brute force and the tutorial known answers.  Prints
// "shim ok" and exits 0 on success.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <random>
#include <vector>

#include "frovedis_hip/join.hpp"

#define EXPECT(c)                                                    \
  do {                                                               \
    if (!(c)) {                                                      \
      std::fprintf(stderr, "FAILED %s at line %d\n", #c, __LINE__); \
      std::exit(1);                                                  \
    }                                                                \
  } while (0)

int main() {
  std::mt19937_64 gen(7);
  // radix_sort(vector<K>&, vector<V>&): equals std::stable_sort by key
  std::vector<long> k(100000);
  for (auto& x : k) x = long(gen() % 5000) - 2500;
  std::vector<size_t> v(k.size());
  std::iota(v.begin(), v.end(), 0);
  std::vector<size_t> order(v);
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return k[a] < k[b]; });
  std::vector<long> ks(k);
  frovedis::radix_sort(ks, v);
  for (size_t i = 0; i < k.size(); i++) EXPECT(v[i] == order[i] && ks[i] == k[order[i]]);
  // size mismatch throws like the reference (radix_sort.hpp:493-494)
  bool threw = false;
  try {
    std::vector<size_t> bad(3);
    frovedis::radix_sort(ks, bad);
  } catch (std::runtime_error&) {
    threw = true;
  }
  EXPECT(threw);
  // set_separate
  auto sep = frovedis::set_separate(ks);
  EXPECT(sep.front() == 0 && sep.back() == ks.size());
  for (size_t i = 1; i + 1 < sep.size(); i++) EXPECT(ks[sep[i]] != ks[sep[i] - 1]);
  EXPECT(frovedis::set_separate(std::vector<int>()).size() == 1);
  // set_is_unique through the device kernel
  EXPECT(frovedis::set_is_unique(ks) == (sep.size() == ks.size() + 1));
  std::vector<long> inc(3000);
  std::iota(inc.begin(), inc.end(), -1000);
  EXPECT(frovedis::set_is_unique(inc) == 1);
  inc[2999] = inc[2998];
  EXPECT(frovedis::set_is_unique(inc) == 0);
  EXPECT(frovedis::set_is_unique(std::vector<int>()) == 1);
  // unique_hashtable with is_unique_ok
  std::vector<long> hk(5000);
  for (size_t i = 0; i < hk.size(); i++) hk[i] = long(i) * 7919 % 100003;
  std::vector<size_t> hv(hk.size());
  std::iota(hv.begin(), hv.end(), 100);
  int uniq = 0;
  frovedis::unique_hashtable<long, size_t> ht(hk, hv, uniq);
  EXPECT(uniq == 1);
  std::vector<long> q = {hk[10], -5, hk[4999], 100004};
  std::vector<size_t> miss;
  auto got = ht.lookup(q, miss);
  EXPECT(got[0] == 110 && got[2] == 5099);
  EXPECT(miss.size() == 2 && miss[0] == 1 && miss[1] == 3);
  auto ex = ht.check_existence(q);
  EXPECT(ex[0] == 1 && ex[1] == 0 && ex[2] == 1 && ex[3] == 0);
  EXPECT(ht.all_keys().size() == hk.size());
  // tutorial join (tutorial.md:2355-2375): 10 pairs
  std::vector<int> t1c3 = {3, 2, 1, 2, 3, 2, 4, 1}, t2c1 = {1, 2, 3, 2};
  std::vector<size_t> li = {0, 1, 2, 3, 4, 5, 6, 7}, ri = {0, 1, 2, 3}, lo, ro;
  frovedis::equi_join(t1c3, li, t2c1, ri, lo, ro);
  EXPECT(lo.size() == 10);
  // the right keys repeat: the sort-merge path left both inputs stable-sorted
  // with their indices (join.hpp:320-322)
  EXPECT((t1c3 == std::vector<int>{1, 1, 2, 2, 2, 3, 3, 4}));
  EXPECT((li == std::vector<size_t>{2, 7, 1, 3, 5, 0, 4, 6}));
  EXPECT((t2c1 == std::vector<int>{1, 2, 2, 3}));
  EXPECT((ri == std::vector<size_t>{0, 1, 3, 2}));
  // tutorial group_by (tutorial.md:2500-2531)
  std::vector<int> c1 = {1, 3, 3, 2, 2, 2, 1, 1}, c3 = {30, 20, 30, 20, 30, 40, 40, 50};
  std::vector<size_t> idx(8), split, nulls;
  std::iota(idx.begin(), idx.end(), 0);
  auto keys = frovedis::local_group_by(c1, idx, split);
  EXPECT(keys.size() == 3 && keys[0] == 1 && keys[2] == 3);
  std::vector<std::vector<size_t>> hd(1, std::vector<size_t>{0, 1, 2});
  auto sums = frovedis::sum_helper(c3, idx, split, hd, nulls);
  auto maxs = frovedis::max_helper(c3, idx, split, hd, nulls);
  auto cnts = frovedis::count_helper(c3, idx, split, hd, nulls);
  EXPECT(sums[0][0] == 120 && sums[0][1] == 90 && sums[0][2] == 50);
  EXPECT(maxs[0][0] == 50 && maxs[0][1] == 40 && maxs[0][2] == 30);
  EXPECT(cnts[0][0] == 3 && cnts[0][1] == 3 && cnts[0][2] == 2);
  std::printf("shim ok\n");
  return 0;
}
