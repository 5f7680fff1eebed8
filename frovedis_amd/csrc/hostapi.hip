// Host-pointer entry points (the reference's std::vector storage) over the
// device entry points: inputs are copied to the device through pinned
// staging, the device path runs on the default stream, results are copied
// back the same way.
#include "fhip_common.hpp"
#include "hashtable.hpp"

extern "C" {
int fhip_dev_ht_lookup(const fhip_ht* ht, const void* keys, size_t n, uint64_t* vals_out,
                       size_t* miss_out, size_t* miss_n, void* stream);
int fhip_dev_ht_check_existence(const fhip_ht* ht, const void* keys, size_t n, int* out,
                                void* stream);
int fhip_dev_ht_all_keys(const fhip_ht* ht, void* out, size_t* out_n, void* stream);
int fhip_dev_group_by(int ktype, const void* col, size_t* idx, size_t n, size_t* split,
                      size_t* nsplit, void* keys_out, void* stream);
int fhip_dev_grouped_agg(int op, int vtype, const void* col, size_t col_n, const size_t* gidx,
                         size_t n, const size_t* split, size_t nsplit, const size_t* nulls,
                         size_t nnull, void* out, void* stream);
}

namespace fhip {
size_t equi_join(int ktype, const void* left, const size_t* lidx, size_t nl, const void* right,
                 const size_t* ridx, size_t nr, int strategy, size_t* lo, size_t* ro,
                 size_t capacity, int* used_hash, size_t* only_left, size_t* n_only,
                 hipStream_t st, bool inplace);
namespace {

// the host entry points run on the default stream, ordered after the
// previous call's stream, and copy through the pinned staging chunks
DBuf to_dev(const void* h, size_t bytes) {
  DBuf d(bytes);
  copy_h2d(d.get(), h, bytes, as_stream(nullptr));
  return d;
}

void to_host(void* h, const void* d, size_t bytes) { copy_d2h(h, d, bytes, as_stream(nullptr)); }

void rethrow(int rc) {
  if (rc != FHIP_OK) throw Error(rc, fhip_last_error());
}

}  // namespace
}  // namespace fhip

extern "C" {

int fhip_ht_build(int ktype, const void* keys, const uint64_t* vals, size_t n, int check_unique,
                  int* is_unique_ok, fhip_ht** ht) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(ht && (n == 0 || keys), FHIP_ERR_ARG, "ht_build: null pointer");
    const size_t kb = size_t(fhip::type_bytes(ktype)) * n;
    fhip::DBuf dk = fhip::to_dev(keys, kb);
    fhip::DBuf dv = vals ? fhip::to_dev(vals, n * 8) : fhip::DBuf();
    *ht = fhip::ht_build(ktype, dk.get(), vals ? dv.as<uint64_t>() : nullptr, n,
                         check_unique != 0, is_unique_ok, nullptr);
    FHIP_CHECK(hipDeviceSynchronize());
  });
}

int fhip_ht_lookup(const fhip_ht* ht, const void* keys, size_t n, uint64_t* vals_out,
                   size_t* miss_out, size_t* miss_n) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(ht && vals_out, FHIP_ERR_ARG, "ht_lookup: null pointer");
    const size_t kb = size_t(fhip::type_bytes(ht->ktype)) * n;
    fhip::DBuf dk = fhip::to_dev(keys, kb), dv(n * 8), dm(n * 8);
    size_t nm = 0;
    fhip::rethrow(fhip_dev_ht_lookup(ht, dk.get(), n, dv.as<uint64_t>(),
                                     miss_out ? dm.as<size_t>() : nullptr, &nm, nullptr));
    fhip::to_host(vals_out, dv.get(), n * 8);
    if (miss_out) fhip::to_host(miss_out, dm.get(), nm * 8);
    if (miss_n) *miss_n = nm;
  });
}

int fhip_ht_check_existence(const fhip_ht* ht, const void* keys, size_t n, int* out) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(ht && out, FHIP_ERR_ARG, "ht_check_existence: null pointer");
    const size_t kb = size_t(fhip::type_bytes(ht->ktype)) * n;
    fhip::DBuf dk = fhip::to_dev(keys, kb), dout(n * sizeof(int));
    fhip::rethrow(fhip_dev_ht_check_existence(ht, dk.get(), n, dout.as<int>(), nullptr));
    fhip::to_host(out, dout.get(), n * sizeof(int));
  });
}

int fhip_ht_all_keys(const fhip_ht* ht, void* out, size_t* out_n) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(ht && out && out_n, FHIP_ERR_ARG, "ht_all_keys: null pointer");
    const size_t kb = size_t(fhip::type_bytes(ht->ktype));
    fhip::DBuf dout(std::max<size_t>(ht->n, 1) * kb);
    size_t m = 0;
    fhip::rethrow(fhip_dev_ht_all_keys(ht, dout.get(), &m, nullptr));
    fhip::to_host(out, dout.get(), m * kb);
    *out_n = m;
  });
}

}  // extern "C"

namespace fhip {
namespace {
// The host forms of equi_join: sorted_back non-null (the in-place form)
// writes the sorted inputs back when the sort-merge path ran.
int host_equi_join(int ktype, const void* left, const size_t* left_idx, size_t nl,
                   const void* right, const size_t* right_idx, size_t nr, int strategy,
                   size_t* left_out, size_t* right_out, size_t capacity, size_t* nout,
                   int* used_hash, void* const* sorted_back) {
  return guarded([&] {
    FHIP_REQUIRE(nout, FHIP_ERR_ARG, "equi_join: null nout");
    const int kb = type_bytes(ktype);
    DBuf dl = to_dev(left, nl * kb), dli = to_dev(left_idx, nl * 8);
    DBuf dr = to_dev(right, nr * kb), dri = to_dev(right_idx, nr * 8);
    DBuf dlo(capacity * 8), dro(capacity * 8);
    int uh = 0;
    size_t r = equi_join(ktype, dl.get(), dli.as<size_t>(), nl, dr.get(), dri.as<size_t>(), nr,
                         strategy, dlo.as<size_t>(), dro.as<size_t>(), capacity, &uh, nullptr,
                         nullptr, nullptr, sorted_back != nullptr);
    if (used_hash) *used_hash = uh;
    *nout = r;
    FHIP_REQUIRE(r <= capacity, FHIP_ERR_TOO_LARGE, "equi_join: output capacity too small");
    to_host(left_out, dlo.get(), r * 8);
    to_host(right_out, dro.get(), r * 8);
    if (sorted_back && !uh) {
      to_host(sorted_back[0], dl.get(), nl * kb);
      to_host(sorted_back[1], dli.get(), nl * 8);
      to_host(sorted_back[2], dr.get(), nr * kb);
      to_host(sorted_back[3], dri.get(), nr * 8);
    }
  });
}
}  // namespace
}  // namespace fhip

extern "C" {

int fhip_equi_join(int ktype, const void* left, const size_t* left_idx, size_t nl,
                   const void* right, const size_t* right_idx, size_t nr, int strategy,
                   size_t* left_out, size_t* right_out, size_t capacity, size_t* nout,
                   int* used_hash) {
  return fhip::host_equi_join(ktype, left, left_idx, nl, right, right_idx, nr, strategy, left_out,
                              right_out, capacity, nout, used_hash, nullptr);
}

int fhip_equi_join_inplace(int ktype, void* left, size_t* left_idx, size_t nl, void* right,
                           size_t* right_idx, size_t nr, int strategy, size_t* left_out,
                           size_t* right_out, size_t capacity, size_t* nout, int* used_hash) {
  void* back[4] = {left, left_idx, right, right_idx};
  return fhip::host_equi_join(ktype, left, left_idx, nl, right, right_idx, nr, strategy, left_out,
                              right_out, capacity, nout, used_hash, back);
}

int fhip_group_by(int ktype, const void* col, size_t col_n, size_t* idx, size_t n, size_t* split,
                  size_t* nsplit, void* keys_out) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(split && nsplit, FHIP_ERR_ARG, "group_by: null output");
    const int kb = fhip::type_bytes(ktype);
    fhip::DBuf dc = fhip::to_dev(col, col_n * kb), di = fhip::to_dev(idx, n * 8);
    fhip::DBuf ds((n + 1) * 8), dk(std::max<size_t>(n, 1) * kb);
    fhip::rethrow(fhip_dev_group_by(ktype, dc.get(), di.as<size_t>(), n, ds.as<size_t>(), nsplit,
                                    dk.get(), nullptr));
    fhip::to_host(idx, di.get(), n * 8);
    fhip::to_host(split, ds.get(), *nsplit * 8);
    fhip::to_host(keys_out, dk.get(), (*nsplit - 1) * kb);
  });
}

int fhip_grouped_agg(int op, int vtype, const void* col, size_t col_n, const size_t* gidx,
                     size_t n, const size_t* split, size_t nsplit, const size_t* nulls,
                     size_t nnull, void* out) {
  return fhip::guarded([&] {
    const int vb = fhip::type_bytes(vtype);
    fhip::DBuf dc = fhip::to_dev(col, col_n * vb), dg = fhip::to_dev(gidx, n * 8);
    fhip::DBuf ds = fhip::to_dev(split, nsplit * 8), dn = fhip::to_dev(nulls, nnull * 8);
    const size_t G = nsplit ? nsplit - 1 : 0;
    const size_t ob = (op == FHIP_AGG_COUNT || op == FHIP_AGG_AVG) ? 8 : size_t(vb);
    fhip::DBuf dout(std::max<size_t>(G, 1) * ob);
    fhip::rethrow(fhip_dev_grouped_agg(op, vtype, dc.get(), col_n, dg.as<size_t>(), n, ds.as<size_t>(),
                                       nsplit, nnull ? dn.as<size_t>() : nullptr, nnull,
                                       dout.get(), nullptr));
    fhip::to_host(out, dout.get(), G * ob);
  });
}

}  // extern "C"
