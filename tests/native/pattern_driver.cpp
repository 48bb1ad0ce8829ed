// Driver for an AddressSanitizer / UndefinedBehaviorSanitizer build of the
// host pattern builder (tensormesh_amd/csrc/tm_pattern.cpp, tm_common.cpp).
// Runs every host entry point on small meshes (a split square, a Kuhn cube,
// an element list with a repeated vertex, an empty mesh) and on malformed
// input (node ids out of range, negative sizes, long rows), checks the
// returned status and a few invariants, and prints "ok" at the end.  Any
// sanitizer report aborts the process with a non-zero status.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "tensormesh_amd.h"

static int failures = 0;
#define CHECK(c)                                                        \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      ++failures;                                                       \
    }                                                                   \
  } while (0)

static std::vector<int32_t> square(int n) {
  std::vector<int32_t> c;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      int a = j * (n + 1) + i, b = a + 1, d = a + n + 1, cc = d + 1;
      c.insert(c.end(), {a, b, cc, a, cc, d});
    }
  return c;
}

static std::vector<int32_t> kuhn(int n) {
  std::vector<int32_t> c;
  const int m = n + 1;
  const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i)
        for (auto& p : perm) {
          int v[3] = {i, j, k};
          c.push_back(v[0] + m * (v[1] + m * v[2]));
          for (int s = 0; s < 3; ++s) {
            ++v[p[s]];
            c.push_back(v[0] + m * (v[1] + m * v[2]));
          }
        }
  return c;
}

static void run(const std::vector<int32_t>& conn, int B, int64_t N) {
  const int64_t E = (int64_t)conn.size() / B;
  int64_t nnz = -1;
  CHECK(tm_pattern_nnz(conn.data(), E, B, N, &nnz) == 0);
  std::vector<int32_t> row_ptr(N + 1), col(nnz), seg_ptr(nnz + 1), seg_slot(E * B * B), s2n(E * B * B);
  CHECK(tm_pattern_build(conn.data(), E, B, N, row_ptr.data(), col.data(), seg_ptr.data(), seg_slot.data(),
                         s2n.data()) == 0);
  CHECK(row_ptr[N] == nnz && seg_ptr[nnz] == E * B * B);
  for (int64_t s = 0; s < E * B * B; ++s) {
    const int64_t e = s / (B * B), i = (s / B) % B, j = s % B;
    const int32_t k = s2n[s];
    CHECK(k >= 0 && k < nnz && col[k] == conn[e * B + j] && k >= row_ptr[conn[e * B + i]] &&
          k < row_ptr[conn[e * B + i] + 1]);
  }
  std::vector<int32_t> inc_ptr(N + 1), inc(E * B);
  CHECK(tm_incidence_build(conn.data(), E, B, N, inc_ptr.data(), inc.data()) == 0);
  std::vector<int32_t> plan(nnz + 1);
  int64_t nb = 0;
  CHECK(tm_reduce_plan(seg_ptr.data(), nnz, 64, plan.data(), &nb) == 0);
  CHECK(nb >= 0 && nb <= nnz && plan[nb] == nnz);
  if (B <= 4) {
    std::vector<uint32_t> cp(E * B), mask(nnz);
    const int st = tm_colpos_build(conn.data(), E, B, N, row_ptr.data(), col.data(), inc_ptr.data(), inc.data(),
                                   cp.data(), mask.data());
    CHECK(st == 0);
  }
}

int main() {
  run(square(5), 3, 36);
  run(kuhn(3), 4, 64);
  // a repeated vertex inside one element
  run(std::vector<int32_t>{0, 1, 1, 1, 2, 3}, 3, 4);
  // no elements
  {
    int64_t nnz = -1;
    CHECK(tm_pattern_nnz(nullptr, 0, 3, 5, &nnz) == 0 && nnz == 0);
  }
  // malformed input: status, no crash
  {
    std::vector<int32_t> bad{0, 1, 7};
    int64_t nnz = -1;
    CHECK(tm_pattern_nnz(bad.data(), 1, 3, 3, &nnz) == TM_ERR_ARG);
    std::vector<int32_t> neg{0, -1, 2};
    CHECK(tm_pattern_nnz(neg.data(), 1, 3, 3, &nnz) == TM_ERR_ARG);
    CHECK(tm_pattern_nnz(bad.data(), -1, 3, 3, &nnz) == TM_ERR_ARG);
    CHECK(tm_last_error()[0] != '\0');
  }
  // a row longer than the packed column positions allow
  {
    const int fan = 300;
    std::vector<int32_t> c;
    for (int t = 0; t < fan; ++t) c.insert(c.end(), {0, 1 + t, 2 + t});
    const int64_t N = fan + 2, E = fan;
    int64_t nnz = 0;
    CHECK(tm_pattern_nnz(c.data(), E, 3, N, &nnz) == 0);
    std::vector<int32_t> rp(N + 1), col(nnz), sp(nnz + 1), ss(E * 9), s2n(E * 9), ip(N + 1), inc(E * 3);
    CHECK(tm_pattern_build(c.data(), E, 3, N, rp.data(), col.data(), sp.data(), ss.data(), s2n.data()) == 0);
    CHECK(tm_incidence_build(c.data(), E, 3, N, ip.data(), inc.data()) == 0);
    std::vector<uint32_t> cp(E * 3), mask(nnz);
    CHECK(tm_colpos_build(c.data(), E, 3, N, rp.data(), col.data(), ip.data(), inc.data(), cp.data(), mask.data()) ==
          TM_ERR_SIZE);
  }
  if (failures) return 1;
  std::puts("ok");
  return 0;
}
