// Fixed-order sum of n doubles into one value of the mesh dtype: block p of
// kSumParts sums a contiguous chunk (strided per thread, then an LDS tree)
// and one block sums the partials, so the result does not depend on
// scheduling.  Used for the scalar energies (J2 plasticity, penalty contact).
// Also a zero-fill kernel for outputs that receive no contribution (used in
// place of a memset, so the calls stay safe to capture in a graph).
#include <hip/hip_runtime.h>

#include "tm_internal.h"

namespace {

constexpr int kSumBlock = 256;

__device__ double block_tree_sum(double v, double* s_red) {
  s_red[threadIdx.x] = v;
  __syncthreads();
  for (int off = kSumBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) s_red[threadIdx.x] += s_red[threadIdx.x + off];
    __syncthreads();
  }
  return s_red[0];
}

__global__ void __launch_bounds__(kSumBlock)
    sum_partials_kernel(const double* __restrict__ v, int64_t n, double* __restrict__ part) {
  __shared__ double s_red[kSumBlock];
  const int64_t chunk = (n + tmh::kSumParts - 1) / tmh::kSumParts;
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < n ? lo + chunk : n;
  double s = 0.0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kSumBlock) s += v[i];
  const double tot = block_tree_sum(s, s_red);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

template <typename T>
__global__ void __launch_bounds__(kSumBlock) sum_final_kernel(const double* __restrict__ part, T* __restrict__ out) {
  __shared__ double s_red[kSumBlock];
  double s = 0.0;
  for (int i = threadIdx.x; i < tmh::kSumParts; i += kSumBlock) s += part[i];
  const double tot = block_tree_sum(s, s_red);
  if (threadIdx.x == 0) out[0] = (T)tot;
}

__global__ void __launch_bounds__(kSumBlock) zero_kernel(uint32_t* __restrict__ p, int64_t n_words) {
  const int64_t i = (int64_t)blockIdx.x * kSumBlock + threadIdx.x;
  if (i < n_words) p[i] = 0u;
}

}  // namespace

namespace tmh {

int zero_fill(void* p, int64_t bytes, tm_stream_t stream) {
  if (bytes <= 0) return TM_OK;
  if (!p || bytes % 4 != 0) return fail(TM_ERR_ARG, "zero fill needs a buffer of whole 4-byte words");
  const int64_t words = bytes / 4;
  zero_kernel<<<(unsigned)((words + kSumBlock - 1) / kSumBlock), kSumBlock, 0,
                reinterpret_cast<hipStream_t>(stream)>>>(static_cast<uint32_t*>(p), words);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    set_error("zero fill launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
}

int sum_fixed(const double* v, int64_t n, double* part, int32_t dtype, void* out, tm_stream_t stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  sum_partials_kernel<<<kSumParts, kSumBlock, 0, s>>>(v, n, part);
  if (dtype == TM_F64)
    sum_final_kernel<double><<<1, kSumBlock, 0, s>>>(part, static_cast<double*>(out));
  else
    sum_final_kernel<float><<<1, kSumBlock, 0, s>>>(part, static_cast<float*>(out));
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    set_error("energy sum launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
}

}  // namespace tmh
