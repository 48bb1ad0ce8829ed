// Batched Jacobi-preconditioned conjugate gradients on CSR systems.
//
// Behavioural reference: the torch CG with Jacobi preconditioner of
// tensormesh/sparse/solve/torch_solve.py:54-128 (which synchronises with the
// host every iteration to test convergence).  Here any number of independent
// SPD systems, stacked block-diagonally in one CSR (rows of system s are
// [sys_row[s], sys_row[s+1])), iterate together:
//   * workgroups are system-aligned (blk_sys / blk_row0 / blk_row1), so every
//     dot product is reduced per system from per-workgroup partials in a fixed
//     order -- results are bitwise reproducible;
//   * all scalars (rz, alpha, beta, |r|^2, |b|^2, active flag) live in device
//     memory, so an iteration is five launches with no host round trip, and a
//     chunk of iterations can be captured in one graph;
//   * a system stops updating once |r| <= rtol |b| (its active flag clears),
//     the host only polls between chunks.
// Scalars layout per system s (double): sc[s*8 + {0:rz, 1:pq, 2:alpha, 3:beta,
// 4:rr, 5:bb, 6:active, 7:iterations}].
#include <hip/hip_runtime.h>

#include "tm_internal.h"

namespace {

constexpr int kBlock = 256;

enum { kRZ = 0, kPQ = 1, kAlpha = 2, kBeta = 3, kRR = 4, kBB = 5, kActive = 6, kIter = 7, kNS = 8 };

template <typename T>
__device__ __forceinline__ double block_sum(double v, double* sh) {
  // deterministic tree sum over the workgroup
  sh[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

template <typename T>
__device__ __forceinline__ T row_dot(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                     const T* __restrict__ av, const T* __restrict__ x, int32_t i) {
  T s = T(0);
  for (int32_t k = rp[i]; k < rp[i + 1]; ++k) s += av[k] * x[ci[k]];
  return s;
}

// r = b - A x, z = dinv r, p = z; partials: rz, bb, rr
template <typename T>
__global__ void __launch_bounds__(kBlock)
    pcg_init_kernel(const int32_t* __restrict__ blk_sys, const int32_t* __restrict__ blk_row0,
                    const int32_t* __restrict__ blk_row1, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                    const T* __restrict__ av, const T* __restrict__ b, const T* __restrict__ x,
                    const T* __restrict__ dinv, T* __restrict__ r, T* __restrict__ z, T* __restrict__ p,
                    double* __restrict__ part) {
  __shared__ double sh[kBlock];
  const int32_t blk = blockIdx.x;
  const int32_t i = blk_row0[blk] + threadIdx.x;
  double rz = 0.0, bb = 0.0, rr = 0.0;
  if (i < blk_row1[blk]) {
    const T ri = b[i] - row_dot(rp, ci, av, x, i);
    const T zi = dinv[i] * ri;
    r[i] = ri;
    z[i] = zi;
    p[i] = zi;
    rz = (double)ri * (double)zi;
    bb = (double)b[i] * (double)b[i];
    rr = (double)ri * (double)ri;
  }
  rz = block_sum<T>(rz, sh);
  bb = block_sum<T>(bb, sh);
  rr = block_sum<T>(rr, sh);
  if (threadIdx.x == 0) {
    part[blk * 3 + 0] = rz;
    part[blk * 3 + 1] = bb;
    part[blk * 3 + 2] = rr;
  }
  (void)blk_sys;
}

// q = A p; partial p.q   (skips inactive systems)
template <typename T>
__global__ void __launch_bounds__(kBlock)
    pcg_spmv_kernel(const int32_t* __restrict__ blk_sys, const int32_t* __restrict__ blk_row0,
                    const int32_t* __restrict__ blk_row1, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                    const T* __restrict__ av, const T* __restrict__ p, T* __restrict__ q,
                    const double* __restrict__ sc, double* __restrict__ part) {
  __shared__ double sh[kBlock];
  const int32_t blk = blockIdx.x;
  const int32_t s = blk_sys[blk];
  if (sc[s * kNS + kActive] == 0.0) return;  // uniform per workgroup
  const int32_t i = blk_row0[blk] + threadIdx.x;
  double pq = 0.0;
  if (i < blk_row1[blk]) {
    const T qi = row_dot(rp, ci, av, p, i);
    q[i] = qi;
    pq = (double)p[i] * (double)qi;
  }
  pq = block_sum<T>(pq, sh);
  if (threadIdx.x == 0) part[blk * 3 + 0] = pq;
}

// x += alpha p; r -= alpha q; z = dinv r; partials r.z, r.r
template <typename T>
__global__ void __launch_bounds__(kBlock)
    pcg_update_kernel(const int32_t* __restrict__ blk_sys, const int32_t* __restrict__ blk_row0,
                      const int32_t* __restrict__ blk_row1, const T* __restrict__ p, const T* __restrict__ q,
                      const T* __restrict__ dinv, T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                      const double* __restrict__ sc, double* __restrict__ part) {
  __shared__ double sh[kBlock];
  const int32_t blk = blockIdx.x;
  const int32_t s = blk_sys[blk];
  if (sc[s * kNS + kActive] == 0.0) return;
  const T alpha = (T)sc[s * kNS + kAlpha];
  const int32_t i = blk_row0[blk] + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (i < blk_row1[blk]) {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    const T zi = dinv[i] * ri;
    r[i] = ri;
    z[i] = zi;
    rz = (double)ri * (double)zi;
    rr = (double)ri * (double)ri;
  }
  rz = block_sum<T>(rz, sh);
  rr = block_sum<T>(rr, sh);
  if (threadIdx.x == 0) {
    part[blk * 3 + 0] = rz;
    part[blk * 3 + 2] = rr;
  }
}

// p = z + beta p
template <typename T>
__global__ void __launch_bounds__(kBlock)
    pcg_dir_kernel(const int32_t* __restrict__ blk_sys, const int32_t* __restrict__ blk_row0,
                   const int32_t* __restrict__ blk_row1, const T* __restrict__ z, T* __restrict__ p,
                   const double* __restrict__ sc) {
  const int32_t blk = blockIdx.x;
  const int32_t s = blk_sys[blk];
  if (sc[s * kNS + kActive] == 0.0) return;
  const T beta = (T)sc[s * kNS + kBeta];
  const int32_t i = blk_row0[blk] + threadIdx.x;
  if (i < blk_row1[blk]) p[i] = z[i] + beta * p[i];
}

// Per-system reductions of the workgroup partials: one workgroup per system,
// thread t adds partials t, t+256, ... in order, then a fixed LDS tree -- the
// same order every run.
//   stage 0 (after init):   rz, bb, rr; active = rr > tol^2 bb
//   stage 1 (after spmv):   pq; alpha = rz / pq
//   stage 2 (after update): rz_new, rr; beta = rz_new / rz; rz = rz_new;
//                           iterations += 1; active = rr > tol^2 bb
__global__ void __launch_bounds__(kBlock)
    pcg_scalar_kernel(int32_t n_sys, const int32_t* __restrict__ sys_blk, const double* __restrict__ part,
                      double* __restrict__ sc, double tol2, int stage) {
  __shared__ double sh[3][kBlock];
  const int32_t s = blockIdx.x;
  if (s >= n_sys) return;
  double* c = sc + s * kNS;
  if (stage != 0 && c[kActive] == 0.0) return;  // uniform per workgroup
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int32_t b = sys_blk[s] + threadIdx.x; b < sys_blk[s + 1]; b += kBlock) {
    a0 += part[b * 3 + 0];
    a1 += part[b * 3 + 1];
    a2 += part[b * 3 + 2];
  }
  sh[0][threadIdx.x] = a0;
  sh[1][threadIdx.x] = a1;
  sh[2][threadIdx.x] = a2;
  __syncthreads();
#pragma unroll
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
      sh[2][threadIdx.x] += sh[2][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  a0 = sh[0][0];
  a1 = sh[1][0];
  a2 = sh[2][0];
  if (stage == 0) {
    c[kRZ] = a0;
    c[kBB] = a1;
    c[kRR] = a2;
    c[kIter] = 0.0;
    c[kActive] = (a2 > tol2 * a1 && a1 > 0.0) ? 1.0 : 0.0;
  } else if (stage == 1) {
    c[kPQ] = a0;
    c[kAlpha] = a0 != 0.0 ? c[kRZ] / a0 : 0.0;
  } else {
    c[kBeta] = c[kRZ] != 0.0 ? a0 / c[kRZ] : 0.0;
    c[kRZ] = a0;
    c[kRR] = a2;
    c[kIter] += 1.0;
    c[kActive] = a2 > tol2 * c[kBB] ? 1.0 : 0.0;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    spmv_kernel(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const T* __restrict__ av,
                const T* __restrict__ x, T* __restrict__ y, int64_t n_rows) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n_rows) return;
  y[i] = row_dot(rp, ci, av, x, (int32_t)i);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    jacobi_kernel(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const T* __restrict__ av,
                  T* __restrict__ dinv, int64_t n_rows) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n_rows) return;
  T d = T(0);
  for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
    if (ci[k] == (int32_t)i) d += av[k];
  dinv[i] = d != T(0) ? T(1) / d : T(1);
}

int hip_status(const char* what) {
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("%s: %s", what, hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
}

template <typename T>
int pcg_init(const tm_pcg_view* v, const void* b, const void* x, hipStream_t s) {
  const int64_t nb = v->n_blocks;
  pcg_init_kernel<T><<<nb, kBlock, 0, s>>>(v->blk_sys, v->blk_row0, v->blk_row1, v->row_ptr, v->col_idx,
                                           (const T*)v->values, (const T*)b, (const T*)x, (const T*)v->dinv,
                                           (T*)v->r, (T*)v->z, (T*)v->p, v->partials);
  pcg_scalar_kernel<<<v->n_sys, kBlock, 0, s>>>((int32_t)v->n_sys, v->sys_blk, v->partials, v->scalars,
                                                v->rtol * v->rtol, 0);
  return hip_status("pcg init");
}

template <typename T>
int pcg_iterate(const tm_pcg_view* v, void* x, int32_t n_iter, hipStream_t s) {
  const int64_t nb = v->n_blocks;
  const int64_t gs = v->n_sys;
  const double tol2 = v->rtol * v->rtol;
  for (int32_t it = 0; it < n_iter; ++it) {
    pcg_spmv_kernel<T><<<nb, kBlock, 0, s>>>(v->blk_sys, v->blk_row0, v->blk_row1, v->row_ptr, v->col_idx,
                                             (const T*)v->values, (const T*)v->p, (T*)v->q, v->scalars, v->partials);
    pcg_scalar_kernel<<<gs, kBlock, 0, s>>>((int32_t)v->n_sys, v->sys_blk, v->partials, v->scalars, tol2, 1);
    pcg_update_kernel<T><<<nb, kBlock, 0, s>>>(v->blk_sys, v->blk_row0, v->blk_row1, (const T*)v->p, (const T*)v->q,
                                               (const T*)v->dinv, (T*)x, (T*)v->r, (T*)v->z, v->scalars, v->partials);
    pcg_scalar_kernel<<<gs, kBlock, 0, s>>>((int32_t)v->n_sys, v->sys_blk, v->partials, v->scalars, tol2, 2);
    pcg_dir_kernel<T><<<nb, kBlock, 0, s>>>(v->blk_sys, v->blk_row0, v->blk_row1, (const T*)v->z, (T*)v->p,
                                            v->scalars);
  }
  return hip_status("pcg iterate");
}

int check_view(const tm_pcg_view* v) {
  if (!v) return tmh::fail(TM_ERR_ARG, "NULL view");
  if (v->n_rows < 0 || v->n_rows >= tmh::kMaxIndex || v->n_blocks <= 0 || v->n_sys <= 0)
    return tmh::fail(TM_ERR_SIZE, "bad system sizes");
  if (!v->row_ptr || !v->col_idx || !v->values || !v->dinv || !v->r || !v->z || !v->p || !v->q || !v->partials ||
      !v->scalars || !v->blk_sys || !v->blk_row0 || !v->blk_row1 || !v->sys_blk)
    return tmh::fail(TM_ERR_ARG, "NULL buffer in pcg view");
  return TM_OK;
}

}  // namespace

extern "C" int tm_spmv(int32_t dtype, int64_t n_rows, const int32_t* row_ptr, const int32_t* col_idx,
                       const void* values, const void* x, void* y, tm_stream_t stream_) {
  TM_TRY_BEGIN
  hipStream_t s = reinterpret_cast<hipStream_t>(stream_);
  if (n_rows < 0 || n_rows >= tmh::kMaxIndex) return tmh::fail(TM_ERR_SIZE, "n_rows out of range");
  if (n_rows == 0) return TM_OK;
  if (!row_ptr || !col_idx || !values || !x || !y) return tmh::fail(TM_ERR_ARG, "NULL pointer");
  const int64_t grid = (n_rows + kBlock - 1) / kBlock;
  if (dtype == TM_F64)
    spmv_kernel<double><<<grid, kBlock, 0, s>>>(row_ptr, col_idx, (const double*)values, (const double*)x, (double*)y, n_rows);
  else if (dtype == TM_F32)
    spmv_kernel<float><<<grid, kBlock, 0, s>>>(row_ptr, col_idx, (const float*)values, (const float*)x, (float*)y, n_rows);
  else
    return tmh::fail(TM_ERR_ARG, "unknown dtype");
  return hip_status("spmv");
  TM_TRY_END
}

extern "C" int tm_jacobi(int32_t dtype, int64_t n_rows, const int32_t* row_ptr, const int32_t* col_idx,
                         const void* values, void* dinv, tm_stream_t stream_) {
  TM_TRY_BEGIN
  hipStream_t s = reinterpret_cast<hipStream_t>(stream_);
  if (n_rows < 0 || n_rows >= tmh::kMaxIndex) return tmh::fail(TM_ERR_SIZE, "n_rows out of range");
  if (n_rows == 0) return TM_OK;
  if (!row_ptr || !col_idx || !values || !dinv) return tmh::fail(TM_ERR_ARG, "NULL pointer");
  const int64_t grid = (n_rows + kBlock - 1) / kBlock;
  if (dtype == TM_F64)
    jacobi_kernel<double><<<grid, kBlock, 0, s>>>(row_ptr, col_idx, (const double*)values, (double*)dinv, n_rows);
  else if (dtype == TM_F32)
    jacobi_kernel<float><<<grid, kBlock, 0, s>>>(row_ptr, col_idx, (const float*)values, (float*)dinv, n_rows);
  else
    return tmh::fail(TM_ERR_ARG, "unknown dtype");
  return hip_status("jacobi");
  TM_TRY_END
}

extern "C" int tm_pcg_init(const tm_pcg_view* view, const void* b, const void* x, tm_stream_t stream_) {
  TM_TRY_BEGIN
  int st = check_view(view);
  if (st) return st;
  if (!b || !x) return tmh::fail(TM_ERR_ARG, "NULL b / x");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream_);
  if (view->dtype == TM_F64) return pcg_init<double>(view, b, x, s);
  if (view->dtype == TM_F32) return pcg_init<float>(view, b, x, s);
  return tmh::fail(TM_ERR_ARG, "unknown dtype");
  TM_TRY_END
}

extern "C" int tm_pcg_iterate(const tm_pcg_view* view, void* x, int32_t n_iter, tm_stream_t stream_) {
  TM_TRY_BEGIN
  int st = check_view(view);
  if (st) return st;
  if (!x || n_iter < 0) return tmh::fail(TM_ERR_ARG, "bad x / n_iter");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream_);
  if (view->dtype == TM_F64) return pcg_iterate<double>(view, x, n_iter, s);
  if (view->dtype == TM_F32) return pcg_iterate<float>(view, x, n_iter, s);
  return tmh::fail(TM_ERR_ARG, "unknown dtype");
  TM_TRY_END
}
