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

// Every cell type as X(enum, D, B), for switch statements over a cell code.
#define TM_SIMPLEX_CELLS(X) \
  X(TM_TRI3, 2, 3) X(TM_TET4, 3, 4) X(TM_TRI6, 2, 6) X(TM_TET10, 3, 10) X(TM_TRI10, 2, 10) X(TM_TET20, 3, 20)
#define TM_TENSOR_CELLS(X) X(TM_QUAD4, 2, 4) X(TM_QUAD9, 2, 9) X(TM_HEX8, 3, 8) X(TM_HEX27, 3, 27)
#define TM_ALL_CELLS(X) TM_SIMPLEX_CELLS(X) TM_TENSOR_CELLS(X)

inline int cell_basis(int cell) {
  switch (cell) {
#define TM_CB(C, D, B) case C: return B;
    TM_ALL_CELLS(TM_CB)
#undef TM_CB
    default: return -1;
  }
}

inline int cell_dim(int cell) {
  switch (cell) {
#define TM_CD(C, D, B) case C: return D;
    TM_ALL_CELLS(TM_CD)
#undef TM_CD
    default: return -1;
  }
}

// Facets per cell: triangle 3 edges, tetrahedron 4 faces, quadrilateral 4
// edges, hexahedron 6 faces.
inline int cell_facets(int cell) {
  switch (cell) {
    case TM_TRI3: case TM_TRI6: case TM_TRI10: return 3;
    case TM_TET4: case TM_TET10: case TM_TET20: case TM_QUAD4: case TM_QUAD9: return 4;
    case TM_HEX8: case TM_HEX27: return 6;
    default: return -1;
  }
}

constexpr int64_t kMaxIndex = 2147483647LL;

// Fixed-order sum of v[0, n) into out[0] (mesh dtype); part: kSumParts doubles.
constexpr int kSumParts = 1024;
int sum_fixed(const double* v, int64_t n, double* part, int32_t dtype, void* out, tm_stream_t stream);
// Zeroes `bytes` (a multiple of 4) of device memory with a kernel.
int zero_fill(void* p, int64_t bytes, tm_stream_t stream);
// Segmented reduction of a pattern view (paired outputs when pat->entry is set).
int reduce_view(int32_t dtype, const tm_pattern_view* pat, int32_t hh, const void* slot, void* out,
                tm_stream_t stream);

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
