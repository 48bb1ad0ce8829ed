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

// ---------------------------------------------------------------------------
// Two-launch PCG iteration (Chronopoulos-Gear recurrences): per iteration one
// elementwise kernel (p, s, x, r, u updates and gamma = (r, u), |r|^2) and
// one SpMV kernel (w = A u and delta = (w, u)); the per-system sums are closed
// by the last workgroup of each system to finish (atomic ticket), which adds
// the partials in workgroup order -- deterministic -- and publishes the next
// alpha / beta.  Recurrences (u = M r, w = A u, s = A p):
//   p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s, u = M r
//   beta' = gamma' / gamma,  alpha' = gamma' / (delta' - beta' gamma' / alpha)
// Extra scalars (after the 8 of the five-launch variant are reused):
//   sc[s*8 + 0] gamma, 1 delta, 2 alpha, 3 beta, 4 |r|^2, 5 |b|^2, 6 active, 7 iterations.
// ---------------------------------------------------------------------------

// Sums part[b*3 + c] over the system's workgroups in order (only the last
// workgroup of a system calls it); loads bypass the non-coherent L1.
__device__ __forceinline__ double system_sum(const double* part, int32_t b0, int32_t b1, int c, double* sh) {
  double v = 0.0;
  for (int32_t b = b0 + threadIdx.x; b < b1; b += kBlock)
    v += __hip_atomic_load(&part[b * 3 + c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return block_sum<double>(v, sh);
}

// Ticket for "am I the last workgroup of system s"; resets the counter.
__device__ __forceinline__ bool last_block(int32_t* counter, int32_t n_blocks_sys, int* sh_flag) {
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const int32_t t = atomicAdd(counter, 1);
    *sh_flag = (t == n_blocks_sys - 1) ? 1 : 0;
    if (*sh_flag) atomicExch(counter, 0);
  }
  __syncthreads();
  const bool last = *sh_flag != 0;
  if (last) __threadfence();
  return last;
}

// r = b - A x, u = dinv r, p = s = 0; gamma = (r, u), |b|^2, |r|^2
template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg2_init_kernel(const int32_t* __restrict__ blk_sys, const int32_t* __restrict__ blk_row0,
                    const int32_t* __restrict__ blk_row1, const int32_t* __restrict__ sys_blk,
                    const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const T* __restrict__ av,
                    const T* __restrict__ b, const T* __restrict__ x, const T* __restrict__ dinv, T* __restrict__ r,
                    T* __restrict__ u, T* __restrict__ p, T* __restrict__ sv, double* __restrict__ part,
                    double* __restrict__ sc, int32_t* __restrict__ counters, double tol2) {
  __shared__ double sh[kBlock];
  __shared__ int flag;
  const int32_t blk = blockIdx.x;
  const int32_t sy = blk_sys[blk];
  const int32_t i = blk_row0[blk] + threadIdx.x;
  double g = 0.0, bb = 0.0, rr = 0.0;
  if (i < blk_row1[blk]) {
    const T ri = b[i] - row_dot(rp, ci, av, x, i);
    const T ui = dinv[i] * ri;
    r[i] = ri;
    u[i] = ui;
    p[i] = T(0);
    sv[i] = T(0);
    g = (double)ri * (double)ui;
    bb = (double)b[i] * (double)b[i];
    rr = (double)ri * (double)ri;
  }
  g = block_sum<T>(g, sh);
  bb = block_sum<T>(bb, sh);
  rr = block_sum<T>(rr, sh);
  if (threadIdx.x == 0) {
    part[blk * 3 + 0] = g;
    part[blk * 3 + 1] = bb;
    part[blk * 3 + 2] = rr;
  }
  const int32_t b0 = sys_blk[sy], b1 = sys_blk[sy + 1];
  if (!last_block(&counters[sy], b1 - b0, &flag)) return;
  const double G = system_sum(part, b0, b1, 0, sh);
  const double BB = system_sum(part, b0, b1, 1, sh);
  const double RR = system_sum(part, b0, b1, 2, sh);
  if (threadIdx.x == 0) {
    double* c = sc + sy * kNS;
    c[kRZ] = G;
    c[kBB] = BB;
    c[kRR] = RR;
    c[kBeta] = 0.0;
    c[kIter] = 0.0;
    c[kActive] = (RR > tol2 * BB && BB > 0.0) ? 1.0 : 0.0;
  }
}

// w = A u; delta = (w, u); last workgroup: alpha (first=1: alpha = gamma / delta)
template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg2_spmv_kernel(const int32_t* __restrict__ blk_sys, const int32_t* __restrict__ blk_row0,
                    const int32_t* __restrict__ blk_row1, const int32_t* __restrict__ sys_blk,
                    const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const T* __restrict__ av,
                    const T* __restrict__ u, T* __restrict__ w, double* __restrict__ part, double* __restrict__ sc,
                    int32_t* __restrict__ counters, int first) {
  __shared__ double sh[kBlock];
  __shared__ int flag;
  const int32_t blk = blockIdx.x;
  const int32_t sy = blk_sys[blk];
  if (sc[sy * kNS + kActive] == 0.0) return;  // uniform per workgroup and per system
  const int32_t i = blk_row0[blk] + threadIdx.x;
  double d = 0.0;
  if (i < blk_row1[blk]) {
    const T wi = row_dot(rp, ci, av, u, i);
    w[i] = wi;
    d = (double)wi * (double)u[i];
  }
  d = block_sum<T>(d, sh);
  if (threadIdx.x == 0) part[blk * 3 + 0] = d;
  const int32_t b0 = sys_blk[sy], b1 = sys_blk[sy + 1];
  if (!last_block(&counters[sy], b1 - b0, &flag)) return;
  const double D = system_sum(part, b0, b1, 0, sh);
  if (threadIdx.x == 0) {
    double* c = sc + sy * kNS;
    const double g = c[kRZ];
    double den = first ? D : D - c[kBeta] * g / c[kAlpha];
    c[kPQ] = D;
    c[kAlpha] = den != 0.0 ? g / den : 0.0;
  }
}

// p = u + beta p; s = w + beta s; x += alpha p; r -= alpha s; u = dinv r;
// gamma' = (r, u), |r|^2; last workgroup: beta' = gamma'/gamma, convergence.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg2_vec_kernel(const int32_t* __restrict__ blk_sys, const int32_t* __restrict__ blk_row0,
                   const int32_t* __restrict__ blk_row1, const int32_t* __restrict__ sys_blk,
                   const T* __restrict__ dinv, const T* __restrict__ w, T* __restrict__ x, T* __restrict__ r,
                   T* __restrict__ u, T* __restrict__ p, T* __restrict__ sv, double* __restrict__ part,
                   double* __restrict__ sc, int32_t* __restrict__ counters, double tol2) {
  __shared__ double sh[kBlock];
  __shared__ int flag;
  const int32_t blk = blockIdx.x;
  const int32_t sy = blk_sys[blk];
  const double* c0 = sc + sy * kNS;
  if (c0[kActive] == 0.0) return;
  const T alpha = (T)c0[kAlpha], beta = (T)c0[kBeta];
  const int32_t i = blk_row0[blk] + threadIdx.x;
  double g = 0.0, rr = 0.0;
  if (i < blk_row1[blk]) {
    const T pi = u[i] + beta * p[i];
    const T si = w[i] + beta * sv[i];
    p[i] = pi;
    sv[i] = si;
    x[i] += alpha * pi;
    const T ri = r[i] - alpha * si;
    const T ui = dinv[i] * ri;
    r[i] = ri;
    u[i] = ui;
    g = (double)ri * (double)ui;
    rr = (double)ri * (double)ri;
  }
  g = block_sum<T>(g, sh);
  rr = block_sum<T>(rr, sh);
  if (threadIdx.x == 0) {
    part[blk * 3 + 0] = g;
    part[blk * 3 + 2] = rr;
  }
  const int32_t b0 = sys_blk[sy], b1 = sys_blk[sy + 1];
  if (!last_block(&counters[sy], b1 - b0, &flag)) return;
  const double G = system_sum(part, b0, b1, 0, sh);
  const double RR = system_sum(part, b0, b1, 2, sh);
  if (threadIdx.x == 0) {
    double* c = sc + sy * kNS;
    c[kBeta] = c[kRZ] != 0.0 ? G / c[kRZ] : 0.0;
    c[kRZ] = G;
    c[kRR] = RR;
    c[kIter] += 1.0;
    c[kActive] = RR > tol2 * c[kBB] ? 1.0 : 0.0;
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

__global__ void fill_counters(int32_t* c, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) c[i] = 0;
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

template <typename T>
int cg2_init(const tm_pcg_view* v, const void* b, const void* x, hipStream_t s) {
  const int64_t nb = v->n_blocks;
  const double tol2 = v->rtol * v->rtol;
  fill_counters<<<(v->n_sys + 255) / 256, 256, 0, s>>>(v->counters, v->n_sys);
  cg2_init_kernel<T><<<nb, kBlock, 0, s>>>(v->blk_sys, v->blk_row0, v->blk_row1, v->sys_blk, v->row_ptr, v->col_idx,
                                           (const T*)v->values, (const T*)b, (const T*)x, (const T*)v->dinv,
                                           (T*)v->r, (T*)v->z, (T*)v->p, (T*)v->s, v->partials, v->scalars,
                                           v->counters, tol2);
  cg2_spmv_kernel<T><<<nb, kBlock, 0, s>>>(v->blk_sys, v->blk_row0, v->blk_row1, v->sys_blk, v->row_ptr, v->col_idx,
                                           (const T*)v->values, (const T*)v->z, (T*)v->q, v->partials, v->scalars,
                                           v->counters, 1);
  return hip_status("pcg init");
}

template <typename T>
int cg2_iterate(const tm_pcg_view* v, void* x, int32_t n_iter, hipStream_t s) {
  const int64_t nb = v->n_blocks;
  const double tol2 = v->rtol * v->rtol;
  for (int32_t it = 0; it < n_iter; ++it) {
    cg2_vec_kernel<T><<<nb, kBlock, 0, s>>>(v->blk_sys, v->blk_row0, v->blk_row1, v->sys_blk, (const T*)v->dinv,
                                            (const T*)v->q, (T*)x, (T*)v->r, (T*)v->z, (T*)v->p, (T*)v->s,
                                            v->partials, v->scalars, v->counters, tol2);
    cg2_spmv_kernel<T><<<nb, kBlock, 0, s>>>(v->blk_sys, v->blk_row0, v->blk_row1, v->sys_blk, v->row_ptr,
                                             v->col_idx, (const T*)v->values, (const T*)v->z, (T*)v->q, v->partials,
                                             v->scalars, v->counters, 0);
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
  const bool two = view->s && view->counters;
  if (view->dtype == TM_F64) return two ? cg2_init<double>(view, b, x, s) : pcg_init<double>(view, b, x, s);
  if (view->dtype == TM_F32) return two ? cg2_init<float>(view, b, x, s) : pcg_init<float>(view, b, x, s);
  return tmh::fail(TM_ERR_ARG, "unknown dtype");
  TM_TRY_END
}

extern "C" int tm_pcg_iterate(const tm_pcg_view* view, void* x, int32_t n_iter, tm_stream_t stream_) {
  TM_TRY_BEGIN
  int st = check_view(view);
  if (st) return st;
  if (!x || n_iter < 0) return tmh::fail(TM_ERR_ARG, "bad x / n_iter");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream_);
  const bool two = view->s && view->counters;
  if (view->dtype == TM_F64) return two ? cg2_iterate<double>(view, x, n_iter, s) : pcg_iterate<double>(view, x, n_iter, s);
  if (view->dtype == TM_F32) return two ? cg2_iterate<float>(view, x, n_iter, s) : pcg_iterate<float>(view, x, n_iter, s);
  return tmh::fail(TM_ERR_ARG, "unknown dtype");
  TM_TRY_END
}
