// Internal helpers shared by the libtmhip translation units.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/tensormesh_amd.h"

namespace tmh {

// Thread-local last-error message behind tm_last_error().
void set_error(const char* fmt, ...);

inline int fail(int status, const char* what) {
  set_error("%s", what);
  return status;
}

inline int cell_basis(int cell) {
  switch (cell) {
    case TM_TRI3: return 3;
    case TM_TET4: return 4;
    case TM_TRI6: return 6;
    case TM_TET10: return 10;
    default: return -1;
  }
}

inline int cell_dim(int cell) {
  switch (cell) {
    case TM_TRI3: case TM_TRI6: return 2;
    case TM_TET4: case TM_TET10: return 3;
    default: return -1;
  }
}

constexpr int64_t kMaxIndex = 2147483647LL;

// Fixed-order sum of v[0, n) into out[0] (mesh dtype); part: kSumParts doubles.
constexpr int kSumParts = 1024;
int sum_fixed(const double* v, int64_t n, double* part, int32_t dtype, void* out, tm_stream_t stream);

}  // namespace tmh

#define TM_TRY_BEGIN try {
#define TM_TRY_END                                               \
  }                                                              \
  catch (const std::exception& ex) {                             \
    tmh::set_error("exception: %s", ex.what());                   \
    return TM_ERR_ARG;                                           \
  }                                                              \
  catch (...) {                                                  \
    tmh::set_error("unknown exception");                          \
    return TM_ERR_ARG;                                           \
  }
