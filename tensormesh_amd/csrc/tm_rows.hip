// Fused row-centric assembly for P1 scalar forms (no slot buffer).
//
// Same result as the two-stage path of tm_p1.hip + tm_reduce.hip -- geometry
// (element.py:956-968, transformation.py:1265-1272), form and integration
// (element_assembler.py:190-209,416-438), ReduceProjector (projector.py:95-123)
// -- computed from the CSR rows' side.  For node r, its incidence list
// (e, i) (ascending e, conn[e,i] == r) names every element whose local row i
// lands in CSR row r; colpos packs, per incidence, the position of each
// conn[e,j] inside row r.  A workgroup owns a run of whole rows (a reduction
// plan over inc_ptr):
//   1. one thread per incidence recomputes the element's geometry and its
//      local row K_e[i][0..B) and stages it, with its colpos word, in LDS;
//   2. one thread per CSR entry (r, p) walks row r's incidences in order and
//      adds the staged value whose colpos byte equals p.
// Every entry is therefore summed in ascending element order, as the
// reference's sequential index_add_ does, without atomics; each element is
// evaluated B times (once per vertex row) instead of being written to and
// read back from a B*B slot buffer.
#include <hip/hip_runtime.h>

#include "tm_internal.h"

namespace {

constexpr int kBlock = 256;
constexpr int kCapInc = 512;   // incidences staged per workgroup (matches the host plan cap)
constexpr int kMaxRows = 256;  // rows staged per workgroup
constexpr int kMaxEnt = 2048;  // CSR entries staged per workgroup

struct RowConsts {
  double W;        // sum_q w_q
  double wphi[4];  // sum_q w_q phi_b(q)
  double mref[16]; // sum_q w_q phi_i phi_j
};

__device__ __forceinline__ int64_t xcd_logical(int64_t b, int64_t nb) {
  const int64_t q = nb / 8, r = nb % 8;
  const int64_t x = b % 8, y = b / 8;
  return x < r ? x * (q + 1) + y : r * (q + 1) + (x - r) * q + y;
}

// Row i of the local matrix of element e, in double.
template <typename T, int D, int FORM>
__device__ __forceinline__ void local_row(const int32_t* __restrict__ conn, const T* __restrict__ pts,
                                          const T* __restrict__ coeff, const RowConsts& c, int64_t e, int i,
                                          double (&Kr)[D + 1]) {
  constexpr int B = D + 1;
  int v[B];
  if constexpr (D == 3) {
    const int4 q = reinterpret_cast<const int4*>(conn)[e];
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
  } else {
#pragma unroll
    for (int b = 0; b < B; ++b) v[b] = conn[e * B + b];
  }
  double x[B][D];
#pragma unroll
  for (int b = 0; b < B; ++b)
#pragma unroll
    for (int d = 0; d < D; ++d) x[b][d] = (double)pts[(int64_t)v[b] * D + d];
  double n[D][D];
  double det;
  if constexpr (D == 3) {
    double a[3], bb[3], cc[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      a[d] = x[1][d] - x[0][d];
      bb[d] = x[2][d] - x[0][d];
      cc[d] = x[3][d] - x[0][d];
    }
    n[0][0] = bb[1] * cc[2] - bb[2] * cc[1];
    n[0][1] = bb[2] * cc[0] - bb[0] * cc[2];
    n[0][2] = bb[0] * cc[1] - bb[1] * cc[0];
    n[1][0] = cc[1] * a[2] - cc[2] * a[1];
    n[1][1] = cc[2] * a[0] - cc[0] * a[2];
    n[1][2] = cc[0] * a[1] - cc[1] * a[0];
    n[2][0] = a[1] * bb[2] - a[2] * bb[1];
    n[2][1] = a[2] * bb[0] - a[0] * bb[2];
    n[2][2] = a[0] * bb[1] - a[1] * bb[0];
    det = a[0] * n[0][0] + a[1] * n[0][1] + a[2] * n[0][2];
  } else {
    const double a0 = x[1][0] - x[0][0], a1 = x[1][1] - x[0][1];
    const double b0 = x[2][0] - x[0][0], b1 = x[2][1] - x[0][1];
    n[0][0] = b1;  n[0][1] = -b0;
    n[1][0] = -a1; n[1][1] = a0;
    det = a0 * b1 - a1 * b0;
  }
  const double adet = fabs(det);
  if constexpr (FORM == TM_FORM_MASS) {
#pragma unroll
    for (int j = 0; j < B; ++j) Kr[j] = adet * c.mref[i * B + j];
    return;
  } else {
    double g[B][D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        g[k + 1][d] = n[k][d];
        s += n[k][d];
      }
      g[0][d] = -s;
    }
    double coef = c.W;
    if constexpr (FORM == TM_FORM_DIFFUSION_ELEM) {
      coef *= (double)coeff[e];
    } else if constexpr (FORM == TM_FORM_DIFFUSION_NODE) {
      double kq = 0.0;
#pragma unroll
      for (int b = 0; b < B; ++b) kq += (double)coeff[v[b]] * c.wphi[b];
      coef = kq;
    }
    const double scale = coef / adet;
    double gi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double s = g[0][d];
#pragma unroll
      for (int b = 1; b < B; ++b) s = (i == b) ? g[b][d] : s;
      gi[d] = s;
    }
#pragma unroll
    for (int j = 0; j < B; ++j) {
      double s = 0.0;
#pragma unroll
      for (int d = 0; d < D; ++d) s += gi[d] * g[j][d];
      Kr[j] = scale * s;
    }
  }
}

template <typename T, int D, int FORM>
__global__ void __launch_bounds__(kBlock)
    p1_rows_kernel(const int32_t* __restrict__ conn, const T* __restrict__ pts, const T* __restrict__ coeff,
                   RowConsts c, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ inc_ptr,
                   const int32_t* __restrict__ inc, const uint32_t* __restrict__ colpos,
                   const uint32_t* __restrict__ entry_mask, const int32_t* __restrict__ blk_ptr, int64_t n_blocks,
                   T* __restrict__ values) {
  constexpr int B = D + 1;
  __shared__ T s_val[kCapInc * B];
  __shared__ uint32_t s_pos[kCapInc];
  __shared__ int32_t s_rp[kMaxRows + 1];    // row starts, relative to k0
  __shared__ int32_t s_ip[kMaxRows + 1];    // incidence starts, relative to t0
  __shared__ uint16_t s_erow[kMaxEnt];      // entry -> local row
  const int64_t blk = xcd_logical(blockIdx.x, n_blocks);
  const int32_t r0 = blk_ptr[blk], r1 = blk_ptr[blk + 1];
  const int32_t t0 = inc_ptr[r0], t1 = inc_ptr[r1];
  const int32_t k0 = row_ptr[r0], k1 = row_ptr[r1];
  const int32_t n_inc = t1 - t0, n_rows = r1 - r0, n_ent = k1 - k0;
  if (n_inc <= kCapInc && n_rows <= kMaxRows && n_ent <= kMaxEnt) {
    // phase 1: local rows of every incidence, plus the row / incidence maps
    for (int32_t t = threadIdx.x; t < n_inc; t += kBlock) {
      const int32_t ei = inc[t0 + t];
      double Kr[B];
      local_row<T, D, FORM>(conn, pts, coeff, c, ei / B, ei % B, Kr);
#pragma unroll
      for (int j = 0; j < B; ++j) s_val[t * B + j] = (T)Kr[j];
      s_pos[t] = colpos[t0 + t];
    }
    for (int32_t lr = threadIdx.x; lr <= n_rows; lr += kBlock) {
      s_rp[lr] = row_ptr[r0 + lr] - k0;
      s_ip[lr] = inc_ptr[r0 + lr] - t0;
    }
    __syncthreads();
    for (int32_t lr = threadIdx.x; lr < n_rows; lr += kBlock)
      for (int32_t k = s_rp[lr]; k < s_rp[lr + 1]; ++k) s_erow[k] = (uint16_t)lr;
    __syncthreads();
    // phase 2: one thread per entry, incidences in ascending element order
    for (int32_t kl = threadIdx.x; kl < n_ent; kl += kBlock) {
      const int32_t lr = s_erow[kl];
      const uint32_t p = (uint32_t)(kl - s_rp[lr]);
      const int32_t ta = s_ip[lr], tz = s_ip[lr + 1];
      T acc = T(0);
      if (entry_mask && tz - ta <= 32) {
        // only the incidences that contribute, lowest bit (= lowest element) first
        uint32_t m = entry_mask[k0 + kl];
        while (m) {
          const int32_t t = ta + __builtin_ctz(m);
          m &= m - 1u;
          const uint32_t cp = s_pos[t];
          int jj = 0;
#pragma unroll
          for (int j = 1; j < B; ++j) jj = (((cp >> (8 * j)) & 0xffu) == p) ? j : jj;
          acc += s_val[t * B + jj];
        }
      } else {
        for (int32_t t = ta; t < tz; ++t) {
          const uint32_t cp = s_pos[t];
#pragma unroll
          for (int j = 0; j < B; ++j)
            if (((cp >> (8 * j)) & 0xffu) == p) acc += s_val[t * B + j];
        }
      }
      values[k0 + kl] = acc;
    }
    return;
  }
  // oversized run: every entry walks global memory, recomputing local rows
  for (int32_t k = k0 + threadIdx.x; k < k1; k += kBlock) {
    int32_t lo = r0, hi = r1 - 1;
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (row_ptr[mid] <= k) lo = mid; else hi = mid - 1;
    }
    const int32_t r = lo;
    const uint32_t p = (uint32_t)(k - row_ptr[r]);
    T acc = T(0);
    for (int32_t t = inc_ptr[r]; t < inc_ptr[r + 1]; ++t) {
      const uint32_t cp = colpos[t];
#pragma unroll
      for (int j = 0; j < B; ++j) {
        if (((cp >> (8 * j)) & 0xffu) == p) {
          const int32_t ei = inc[t];
          double Kr[B];
          local_row<T, D, FORM>(conn, pts, coeff, c, ei / B, ei % B, Kr);
          acc += (T)Kr[j];
        }
      }
    }
    values[k] = acc;
  }
}

// Row-centric stage S for P1 slot buffers: the same staging and per-entry walk
// as p1_rows_kernel, but phase 1 reads the element rows K_e[i][0..B) from the
// slot buffer written by stage G (one contiguous B-value row per incidence)
// instead of recomputing them.
template <typename T, int D, bool INC_ORDER>
__global__ void __launch_bounds__(kBlock)
    p1_rows_reduce_kernel(const T* __restrict__ slot, const int32_t* __restrict__ row_ptr,
                          const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc,
                          const uint32_t* __restrict__ colpos, const uint32_t* __restrict__ entry_mask,
                          const int32_t* __restrict__ blk_ptr, int64_t n_blocks, T* __restrict__ values) {
  constexpr int B = D + 1;
  __shared__ T s_val[kCapInc * B];
  __shared__ uint32_t s_pos[kCapInc];
  __shared__ int32_t s_rp[kMaxRows + 1];
  __shared__ int32_t s_ip[kMaxRows + 1];
  __shared__ uint16_t s_erow[kMaxEnt];
  const int64_t blk = xcd_logical(blockIdx.x, n_blocks);
  const int32_t r0 = blk_ptr[blk], r1 = blk_ptr[blk + 1];
  const int32_t t0 = inc_ptr[r0], t1 = inc_ptr[r1];
  const int32_t k0 = row_ptr[r0], k1 = row_ptr[r1];
  const int32_t n_inc = t1 - t0, n_rows = r1 - r0, n_ent = k1 - k0;
  if (n_inc <= kCapInc && n_rows <= kMaxRows && n_ent <= kMaxEnt) {
    for (int32_t t = threadIdx.x; t < n_inc; t += kBlock) {
      // element-ordered slots: row (e, i) starts at (e*B + i)*B; incidence-ordered: at (t0 + t)*B
      const T* src = INC_ORDER ? slot + (int64_t)(t0 + t) * B : slot + (int64_t)inc[t0 + t] * B;
      if constexpr (B == 4 && sizeof(T) == 8) {
        const double2 a = reinterpret_cast<const double2*>(src)[0];
        const double2 b = reinterpret_cast<const double2*>(src)[1];
        s_val[t * B + 0] = a.x; s_val[t * B + 1] = a.y;
        s_val[t * B + 2] = b.x; s_val[t * B + 3] = b.y;
      } else if constexpr (B == 4 && sizeof(T) == 4) {
        const float4 a = reinterpret_cast<const float4*>(src)[0];
        s_val[t * B + 0] = a.x; s_val[t * B + 1] = a.y;
        s_val[t * B + 2] = a.z; s_val[t * B + 3] = a.w;
      } else {
#pragma unroll
        for (int j = 0; j < B; ++j) s_val[t * B + j] = src[j];
      }
      s_pos[t] = colpos[t0 + t];
    }
    for (int32_t lr = threadIdx.x; lr <= n_rows; lr += kBlock) {
      s_rp[lr] = row_ptr[r0 + lr] - k0;
      s_ip[lr] = inc_ptr[r0 + lr] - t0;
    }
    __syncthreads();
    for (int32_t lr = threadIdx.x; lr < n_rows; lr += kBlock)
      for (int32_t k = s_rp[lr]; k < s_rp[lr + 1]; ++k) s_erow[k] = (uint16_t)lr;
    __syncthreads();
    for (int32_t kl = threadIdx.x; kl < n_ent; kl += kBlock) {
      const int32_t lr = s_erow[kl];
      const uint32_t p = (uint32_t)(kl - s_rp[lr]);
      const int32_t ta = s_ip[lr], tz = s_ip[lr + 1];
      T acc = T(0);
      if (entry_mask && tz - ta <= 32) {
        // only the incidences that contribute, lowest bit (= lowest element) first
        uint32_t m = entry_mask[k0 + kl];
        while (m) {
          const int32_t t = ta + __builtin_ctz(m);
          m &= m - 1u;
          const uint32_t cp = s_pos[t];
          int jj = 0;
#pragma unroll
          for (int j = 1; j < B; ++j) jj = (((cp >> (8 * j)) & 0xffu) == p) ? j : jj;
          acc += s_val[t * B + jj];
        }
      } else {
        for (int32_t t = ta; t < tz; ++t) {
          const uint32_t cp = s_pos[t];
#pragma unroll
          for (int j = 0; j < B; ++j)
            if (((cp >> (8 * j)) & 0xffu) == p) acc += s_val[t * B + j];
        }
      }
      values[k0 + kl] = acc;
    }
    return;
  }
  for (int32_t k = k0 + threadIdx.x; k < k1; k += kBlock) {
    int32_t lo = r0, hi = r1 - 1;
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (row_ptr[mid] <= k) lo = mid; else hi = mid - 1;
    }
    const int32_t r = lo;
    const uint32_t p = (uint32_t)(k - row_ptr[r]);
    T acc = T(0);
    for (int32_t t = inc_ptr[r]; t < inc_ptr[r + 1]; ++t) {
      const uint32_t cp = colpos[t];
#pragma unroll
      for (int j = 0; j < B; ++j)
        if (((cp >> (8 * j)) & 0xffu) == p) acc += slot[(INC_ORDER ? (int64_t)t : (int64_t)inc[t]) * B + j];
    }
    values[k] = acc;
  }
}

template <typename T, int D>
int launch(const tm_mesh_view* m, const tm_form_desc* f, const tm_rows_view* rv, void* values, hipStream_t s) {
  constexpr int B = D + 1;
  RowConsts c{};
  if (!f->tab_host || f->n_quad <= 0) return tmh::fail(TM_ERR_ARG, "form tables missing");
  const int Q = f->n_quad;
  for (int q = 0; q < Q; ++q) {
    const double w = f->tab_host[q];
    const double* phi = f->tab_host + Q + q * B;
    c.W += w;
    for (int b = 0; b < B; ++b) c.wphi[b] += w * phi[b];
    for (int i = 0; i < B; ++i)
      for (int j = 0; j < B; ++j) c.mref[i * B + j] += w * phi[i] * phi[j];
  }
  const T* pts = static_cast<const T*>(m->points);
  const T* coeff = static_cast<const T*>(f->coeff);
  T* out = static_cast<T*>(values);
  const int64_t nb = rv->n_blocks;
  if (nb == 0) return TM_OK;
#define TM_ROWS_LAUNCH(FORM)                                                                              \
  p1_rows_kernel<T, D, FORM><<<nb, kBlock, 0, s>>>(m->conn, pts, coeff, c, rv->row_ptr, rv->inc_ptr, rv->inc, \
                                                   rv->colpos, rv->entry_mask, rv->blk_ptr, nb, out)
  switch (f->form) {
    case TM_FORM_LAPLACE: TM_ROWS_LAUNCH(TM_FORM_LAPLACE); break;
    case TM_FORM_DIFFUSION_ELEM:
      if (!coeff) return tmh::fail(TM_ERR_ARG, "diffusion form needs a coefficient");
      TM_ROWS_LAUNCH(TM_FORM_DIFFUSION_ELEM);
      break;
    case TM_FORM_DIFFUSION_NODE:
      if (!coeff) return tmh::fail(TM_ERR_ARG, "diffusion form needs a coefficient");
      TM_ROWS_LAUNCH(TM_FORM_DIFFUSION_NODE);
      break;
    case TM_FORM_MASS: TM_ROWS_LAUNCH(TM_FORM_MASS); break;
    default: return tmh::fail(TM_ERR_UNSUPPORTED, "form not available for the fused P1 path");
  }
#undef TM_ROWS_LAUNCH
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("p1 rows kernel launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
}

}  // namespace

extern "C" int tm_assemble_p1_rows(const tm_mesh_view* mesh, const tm_form_desc* form, const tm_rows_view* rows,
                                   void* values, tm_stream_t stream_) {
  TM_TRY_BEGIN
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  if (!mesh || !form || !rows) return tmh::fail(TM_ERR_ARG, "NULL argument");
  if (mesh->cell != TM_TRI3 && mesh->cell != TM_TET4) return tmh::fail(TM_ERR_UNSUPPORTED, "fused path is P1 only");
  if (mesh->dtype != TM_F32 && mesh->dtype != TM_F64) return tmh::fail(TM_ERR_ARG, "unknown dtype");
  if (mesh->n_elem * 4 >= tmh::kMaxIndex || rows->nnz >= tmh::kMaxIndex) return tmh::fail(TM_ERR_SIZE, "too large");
  if (rows->n_blocks > 0 && (!rows->row_ptr || !rows->inc_ptr || !rows->inc || !rows->colpos || !rows->blk_ptr ||
                             !values || !mesh->conn || !mesh->points))
    return tmh::fail(TM_ERR_ARG, "NULL buffer");
  if (mesh->dtype == TM_F64)
    return mesh->cell == TM_TET4 ? launch<double, 3>(mesh, form, rows, values, stream)
                                 : launch<double, 2>(mesh, form, rows, values, stream);
  return mesh->cell == TM_TET4 ? launch<float, 3>(mesh, form, rows, values, stream)
                               : launch<float, 2>(mesh, form, rows, values, stream);
  TM_TRY_END
}

int tm_rows_reduce_launch(int32_t dtype, int32_t cell, const tm_rows_view* rows, const void* slot, bool inc_order,
                          void* values, hipStream_t stream) {
  if (!rows) return tmh::fail(TM_ERR_ARG, "NULL rows view");
  if (cell != TM_TRI3 && cell != TM_TET4) return tmh::fail(TM_ERR_UNSUPPORTED, "row reduction is P1 only");
  const int64_t nb = rows->n_blocks;
  if (nb == 0) return TM_OK;
  if (!slot || !values || !rows->row_ptr || !rows->inc_ptr || !rows->inc || !rows->colpos || !rows->blk_ptr)
    return tmh::fail(TM_ERR_ARG, "NULL buffer");
#define TM_RR(T, D, IO)                                                                                          \
  p1_rows_reduce_kernel<T, D, IO><<<nb, kBlock, 0, stream>>>((const T*)slot, rows->row_ptr, rows->inc_ptr, rows->inc, \
                                                             rows->colpos, rows->entry_mask, rows->blk_ptr, nb,      \
                                                             (T*)values)
  if (dtype == TM_F64) {
    if (cell == TM_TET4) { if (inc_order) TM_RR(double, 3, true); else TM_RR(double, 3, false); }
    else { if (inc_order) TM_RR(double, 2, true); else TM_RR(double, 2, false); }
  } else if (dtype == TM_F32) {
    if (cell == TM_TET4) { if (inc_order) TM_RR(float, 3, true); else TM_RR(float, 3, false); }
    else { if (inc_order) TM_RR(float, 2, true); else TM_RR(float, 2, false); }
  } else {
    return tmh::fail(TM_ERR_ARG, "unknown dtype");
  }
#undef TM_RR
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("row reduction launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
}

extern "C" int tm_reduce_p1_rows(int32_t dtype, int32_t cell, const tm_rows_view* rows, const void* slot,
                                 int32_t inc_order, void* values, tm_stream_t stream_) {
  TM_TRY_BEGIN
  return tm_rows_reduce_launch(dtype, cell, rows, slot, inc_order != 0, values,
                               reinterpret_cast<hipStream_t>(stream_));
  TM_TRY_END
}
