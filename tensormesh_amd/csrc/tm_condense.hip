// Dirichlet condensation of an assembled matrix and its right-hand side.
//
// Replaces the per-call part of Condenser.__call__
// (tensormesh/operator/condense.py:238-296): K_ii = edata[is_inner_edge]
// (order-preserving mask, :280) and f_i = rhs[inner] - K_io @ u_o (:285-296).
// The masks and the local numbering are topology-only and are computed once
// per pattern by the caller (condense.py:182-220); this kernel does the
// value work in one launch with two kinds of workgroups:
//   gather workgroups, one thread per kept entry (coalesced over t):
//     kii[t]   = values[inner_idx[t]]              t in [0, inner_rp[n_inner])
//   row workgroups, one thread per inner row:
//     f_i[i]   = rhs[inner_dof[i]] - sum_t values[ou_idx[t]] * u_o[ou_col[t]]
//                                                  t in [ou_rp[i], ou_rp[i+1])
// The K_io product is summed in ascending entry order.
#include <hip/hip_runtime.h>

#include "tm_internal.h"

namespace {

constexpr int kBlock = 256;

template <typename T>
__global__ void __launch_bounds__(kBlock)
    condense_kernel(int64_t n_inner, const T* __restrict__ values, const int32_t* __restrict__ inner_rp,
                    const int32_t* __restrict__ inner_idx, const int32_t* __restrict__ ou_rp,
                    const int32_t* __restrict__ ou_idx, const int32_t* __restrict__ ou_col,
                    const int32_t* __restrict__ inner_dof, const T* __restrict__ rhs, const T* __restrict__ u_o,
                    T* __restrict__ kii, T* __restrict__ f_i, int64_t n_kept, int64_t gather_blocks) {
  if ((int64_t)blockIdx.x < gather_blocks) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t < n_kept) kii[t] = values[inner_idx[t]];
    return;
  }
  const int64_t i = ((int64_t)blockIdx.x - gather_blocks) * kBlock + threadIdx.x;
  if (i >= n_inner) return;
  T minus = T(0);
  for (int32_t t = ou_rp[i]; t < ou_rp[i + 1]; ++t) minus += values[ou_idx[t]] * u_o[ou_col[t]];
  f_i[i] = rhs[inner_dof[i]] - minus;
}

}  // namespace

extern "C" int tm_condense(int32_t dtype, int64_t n_inner, int64_t n_kept, const void* values, const int32_t* inner_rp,
                           const int32_t* inner_idx, const int32_t* ou_rp, const int32_t* ou_idx,
                           const int32_t* ou_col, const int32_t* inner_dof, const void* rhs, const void* u_o,
                           void* kii, void* f_i, tm_stream_t stream_) {
  TM_TRY_BEGIN
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  if (n_inner < 0 || n_inner >= tmh::kMaxIndex) return tmh::fail(TM_ERR_SIZE, "n_inner out of range");
  if (n_kept < 0 || n_kept >= tmh::kMaxIndex) return tmh::fail(TM_ERR_SIZE, "n_kept out of range");
  if (n_inner == 0) return TM_OK;
  if (!values || !inner_rp || !inner_idx || !ou_rp || !inner_dof || !rhs || !kii || !f_i)
    return tmh::fail(TM_ERR_ARG, "NULL pointer");
  const int64_t gather_blocks = (n_kept + kBlock - 1) / kBlock;
  const int64_t grid = gather_blocks + (n_inner + kBlock - 1) / kBlock;
  if (dtype == TM_F64) {
    condense_kernel<double><<<grid, kBlock, 0, stream>>>(
        n_inner, (const double*)values, inner_rp, inner_idx, ou_rp, ou_idx, ou_col, inner_dof, (const double*)rhs,
        (const double*)u_o, (double*)kii, (double*)f_i, n_kept, gather_blocks);
  } else if (dtype == TM_F32) {
    condense_kernel<float><<<grid, kBlock, 0, stream>>>(
        n_inner, (const float*)values, inner_rp, inner_idx, ou_rp, ou_idx, ou_col, inner_dof, (const float*)rhs,
        (const float*)u_o, (float*)kii, (float*)f_i, n_kept, gather_blocks);
  } else {
    return tmh::fail(TM_ERR_ARG, "unknown dtype");
  }
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("condense kernel launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
  TM_TRY_END
}
