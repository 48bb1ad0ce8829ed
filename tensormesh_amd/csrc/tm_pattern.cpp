// Host-side sparsity pattern and slot->entry permutation.
//
// Replaces the edge construction of ElementAssembler.from_mesh
// (tensormesh/assemble/element_assembler.py:804-853).  The reference pairs
// every slot s = e*B*B + i*B + j with (conn[e,i], conn[e,j]), de-duplicates the
// pairs with scipy coo->csr->coo (sorted unique = CSR order) and looks up each
// slot's entry id.  Here the same pattern comes out of a per-row bucket pass:
// the node incidence lists (element ids ascending) give, for row r, every
// (col, slot) pair with conn[e,i] == r; sorting the bucket by (col, slot)
// yields the row's unique columns and, per entry, its contributing slots in
// ascending slot order -- the order in which ReduceProjector's sequential
// index_add_ adds them (projector.py:118).
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <thread>
#include <vector>

#include "tm_internal.h"

namespace {

struct Incidence {
  std::vector<int32_t> ptr;  // [n_points + 1]
  std::vector<int32_t> inc;  // [n_elem * B], packed e*B + i, ascending e per node
};

int check_conn(const int32_t* conn, int64_t n_elem, int32_t nb, int64_t n_points) {
  if (!conn && n_elem > 0) return tmh::fail(TM_ERR_ARG, "conn is NULL");
  if (n_elem < 0 || nb <= 0 || n_points < 0) return tmh::fail(TM_ERR_ARG, "negative size");
  if (n_points >= tmh::kMaxIndex || n_elem * nb * nb >= tmh::kMaxIndex)
    return tmh::fail(TM_ERR_SIZE, "pattern exceeds int32 indexing (n_elem*B*B or n_points >= 2^31)");
  const int64_t total = n_elem * nb;
  for (int64_t t = 0; t < total; ++t) {
    if (conn[t] < 0 || conn[t] >= n_points) {
      tmh::set_error("conn[%lld] = %d is outside [0, %lld)", (long long)t, conn[t], (long long)n_points);
      return TM_ERR_ARG;
    }
  }
  return TM_OK;
}

void build_incidence(const int32_t* conn, int64_t n_elem, int32_t nb, int64_t n_points, Incidence& out) {
  out.ptr.assign(n_points + 1, 0);
  const int64_t total = n_elem * nb;
  for (int64_t t = 0; t < total; ++t) out.ptr[conn[t] + 1]++;
  for (int64_t n = 0; n < n_points; ++n) out.ptr[n + 1] += out.ptr[n];
  out.inc.resize(total);
  std::vector<int32_t> cursor(out.ptr.begin(), out.ptr.end() - 1);
  for (int64_t t = 0; t < total; ++t) out.inc[cursor[conn[t]]++] = (int32_t)t;  // t = e*B + i
}

int n_threads_for(int64_t work) {
  unsigned hw = std::thread::hardware_concurrency();
  int n = hw == 0 ? 4 : (int)std::min<unsigned>(hw, 16u);
  if (work < 20000) n = 1;
  return n;
}

template <typename Fn>
void parallel_rows(int64_t n_rows, Fn fn) {
  int nt = n_threads_for(n_rows);
  if (nt <= 1) {
    fn(0, n_rows);
    return;
  }
  std::vector<std::thread> pool;
  int64_t chunk = (n_rows + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    int64_t a = t * chunk, b = std::min<int64_t>(n_rows, a + chunk);
    if (a >= b) break;
    pool.emplace_back([=] { fn(a, b); });
  }
  for (auto& th : pool) th.join();
}

// Sorted (col << 32 | slot) keys of row r into buf.
inline void row_keys(const int32_t* conn, int32_t nb, const Incidence& I, int64_t r,
                     std::vector<uint64_t>& buf) {
  buf.clear();
  for (int32_t t = I.ptr[r]; t < I.ptr[r + 1]; ++t) {
    const int64_t e = I.inc[t] / nb;
    const int64_t i = I.inc[t] % nb;
    const int32_t* ce = conn + e * nb;
    const uint64_t slot0 = (uint64_t)(e * nb * nb + i * nb);
    for (int32_t j = 0; j < nb; ++j) buf.push_back(((uint64_t)(uint32_t)ce[j] << 32) | (slot0 + j));
  }
  std::sort(buf.begin(), buf.end());
}

}  // namespace

extern "C" int tm_pattern_nnz(const int32_t* conn, int64_t n_elem, int32_t n_basis, int64_t n_points,
                              int64_t* nnz_out) {
  TM_TRY_BEGIN
  if (!nnz_out) return tmh::fail(TM_ERR_ARG, "nnz_out is NULL");
  int st = check_conn(conn, n_elem, n_basis, n_points);
  if (st) return st;
  Incidence I;
  build_incidence(conn, n_elem, n_basis, n_points, I);
  std::vector<int64_t> rowlen(n_points, 0);
  parallel_rows(n_points, [&](int64_t a, int64_t b) {
    std::vector<uint64_t> buf;
    for (int64_t r = a; r < b; ++r) {
      row_keys(conn, n_basis, I, r, buf);
      int64_t u = 0;
      for (size_t k = 0; k < buf.size(); ++k)
        if (k == 0 || (buf[k] >> 32) != (buf[k - 1] >> 32)) ++u;
      rowlen[r] = u;
    }
  });
  int64_t nnz = 0;
  for (int64_t r = 0; r < n_points; ++r) nnz += rowlen[r];
  if (nnz >= tmh::kMaxIndex) return tmh::fail(TM_ERR_SIZE, "nnz exceeds int32 indexing");
  *nnz_out = nnz;
  return TM_OK;
  TM_TRY_END
}

extern "C" int tm_pattern_build(const int32_t* conn, int64_t n_elem, int32_t n_basis, int64_t n_points,
                                int32_t* row_ptr, int32_t* col_idx, int32_t* seg_ptr, int32_t* seg_slot,
                                int32_t* slot2nnz) {
  TM_TRY_BEGIN
  if (!row_ptr || !col_idx || !seg_ptr || !seg_slot) return tmh::fail(TM_ERR_ARG, "NULL output");
  int st = check_conn(conn, n_elem, n_basis, n_points);
  if (st) return st;
  const int32_t nb = n_basis;
  Incidence I;
  build_incidence(conn, n_elem, nb, n_points, I);

  // pass 1: unique columns per row
  std::vector<int32_t> rowlen(n_points, 0);
  parallel_rows(n_points, [&](int64_t a, int64_t b) {
    std::vector<uint64_t> buf;
    for (int64_t r = a; r < b; ++r) {
      row_keys(conn, nb, I, r, buf);
      int32_t u = 0;
      for (size_t k = 0; k < buf.size(); ++k)
        if (k == 0 || (buf[k] >> 32) != (buf[k - 1] >> 32)) ++u;
      rowlen[r] = u;
    }
  });
  row_ptr[0] = 0;
  int64_t acc = 0;
  for (int64_t r = 0; r < n_points; ++r) {
    acc += rowlen[r];
    if (acc >= tmh::kMaxIndex) return tmh::fail(TM_ERR_SIZE, "nnz exceeds int32 indexing");
    row_ptr[r + 1] = (int32_t)acc;
  }
  const int64_t nnz = acc;

  // pass 2: columns, segments, slot -> entry
  parallel_rows(n_points, [&](int64_t a, int64_t b) {
    std::vector<uint64_t> buf;
    for (int64_t r = a; r < b; ++r) {
      row_keys(conn, nb, I, r, buf);
      int64_t k = (int64_t)row_ptr[r] - 1;
      int64_t off = (int64_t)I.ptr[r] * nb;  // this row's first contribution
      for (size_t t = 0; t < buf.size(); ++t) {
        const uint32_t col = (uint32_t)(buf[t] >> 32);
        const uint32_t slot = (uint32_t)(buf[t] & 0xffffffffu);
        if (t == 0 || (buf[t] >> 32) != (buf[t - 1] >> 32)) {
          ++k;
          col_idx[k] = (int32_t)col;
          seg_ptr[k] = (int32_t)(off + t);
        }
        seg_slot[off + t] = (int32_t)slot;
        if (slot2nnz) slot2nnz[slot] = (int32_t)k;
      }
    }
  });
  seg_ptr[nnz] = (int32_t)(n_elem * nb * nb);
  return TM_OK;
  TM_TRY_END
}

extern "C" int tm_incidence_build(const int32_t* conn, int64_t n_elem, int32_t n_basis, int64_t n_points,
                                  int32_t* inc_ptr, int32_t* inc) {
  TM_TRY_BEGIN
  if (!inc_ptr || !inc) return tmh::fail(TM_ERR_ARG, "NULL output");
  int st = check_conn(conn, n_elem, n_basis, n_points);
  if (st) return st;
  Incidence I;
  build_incidence(conn, n_elem, n_basis, n_points, I);
  std::memcpy(inc_ptr, I.ptr.data(), sizeof(int32_t) * (n_points + 1));
  std::memcpy(inc, I.inc.data(), sizeof(int32_t) * I.inc.size());
  return TM_OK;
  TM_TRY_END
}

extern "C" int tm_reduce_plan(const int32_t* seg_ptr, int64_t n_entries, int32_t cap, int32_t* blk_ptr_out,
                              int64_t* n_blocks) {
  TM_TRY_BEGIN
  if (!seg_ptr || !blk_ptr_out || !n_blocks) return tmh::fail(TM_ERR_ARG, "NULL argument");
  if (n_entries < 0 || n_entries >= tmh::kMaxIndex || cap <= 0) return tmh::fail(TM_ERR_ARG, "bad size");
  int64_t nb = 0;
  int64_t k = 0;
  blk_ptr_out[0] = 0;
  while (k < n_entries) {
    const int64_t base = seg_ptr[k];
    int64_t e = k + 1;  // an entry longer than cap still gets a block of its own
    while (e < n_entries && (int64_t)seg_ptr[e + 1] - base <= cap) ++e;
    blk_ptr_out[++nb] = (int32_t)e;
    k = e;
  }
  *n_blocks = nb;
  return TM_OK;
  TM_TRY_END
}

extern "C" int tm_colpos_build(const int32_t* conn, int64_t n_elem, int32_t n_basis, int64_t n_points,
                               const int32_t* row_ptr, const int32_t* col_idx, const int32_t* inc_ptr,
                               const int32_t* inc, uint32_t* colpos, uint32_t* entry_mask) {
  TM_TRY_BEGIN
  if (!conn || !row_ptr || !col_idx || !inc_ptr || !inc || !colpos) return tmh::fail(TM_ERR_ARG, "NULL argument");
  if (n_basis > 4) return tmh::fail(TM_ERR_UNSUPPORTED, "colpos packs at most 4 positions");
  for (int64_t r = 0; r < n_points; ++r)
    if (row_ptr[r + 1] - row_ptr[r] > 255) return tmh::fail(TM_ERR_SIZE, "a row has more than 255 entries");
  int bad = 0;
  parallel_rows(n_points, [&](int64_t a, int64_t b) {
    for (int64_t r = a; r < b; ++r) {
      const int32_t* cb = col_idx + row_ptr[r];
      const int32_t* ce = col_idx + row_ptr[r + 1];
      for (int32_t t = inc_ptr[r]; t < inc_ptr[r + 1]; ++t) {
        const int64_t e = inc[t] / n_basis;
        uint32_t word = 0;
        for (int32_t j = 0; j < n_basis; ++j) {
          const int32_t c = conn[e * n_basis + j];
          const int32_t* it = std::lower_bound(cb, ce, c);
          if (it == ce || *it != c) { bad = 1; continue; }
          word |= (uint32_t)(it - cb) << (8 * j);
        }
        colpos[t] = word;
      }
      if (entry_mask) {
        // bit (t - inc_ptr[r]) of entry (r, p): incidence t contributes to it
        const int32_t k0 = row_ptr[r], nk = row_ptr[r + 1] - row_ptr[r];
        for (int32_t p = 0; p < nk; ++p) entry_mask[k0 + p] = 0u;
        const int32_t n_inc = inc_ptr[r + 1] - inc_ptr[r];
        if (n_inc <= 32) {
          for (int32_t t = inc_ptr[r]; t < inc_ptr[r + 1]; ++t)
            for (int32_t j = 0; j < n_basis; ++j)
              entry_mask[k0 + ((colpos[t] >> (8 * j)) & 0xffu)] |= 1u << (t - inc_ptr[r]);
        }
      }
    }
  });
  if (bad) return tmh::fail(TM_ERR_ARG, "pattern does not contain an element pair");
  (void)n_elem;
  return TM_OK;
  TM_TRY_END
}
