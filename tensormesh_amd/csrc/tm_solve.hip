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
//     memory, so an iteration needs no host round trip, and a chunk of
//     iterations can be captured in one graph;
//   * a system stops updating once |r| <= rtol |b| (its active flag clears),
//     the host only polls between chunks;
//   * SpMV rows are summed in ascending entry order (CSR-stream through LDS);
//   * logical workgroup b (rows, partial-sum slot) runs on XCD-contiguous
//     hardware workgroups (tmh::xcd_block) in every kernel, so each XCD keeps
//     one contiguous share of the vectors in its L2 from launch to launch
//     and the SpMV's column gathers stay mostly within that share.
// Two iteration variants: the textbook one, five launches per iteration with
// scalars sc[s*8 + {0:rz, 1:pq, 2:alpha, 3:beta, 4:rr, 5:bb, 6:active,
// 7:iterations}], and a two-launch Chronopoulos-Gear one (described where its
// kernels are defined).
#include <hip/hip_runtime.h>

#include "tm_device.h"
#include "tm_internal.h"

// No implicit contraction of a * b + c in this file: a rounding step should
// not depend on how the compiler schedules an expression, so that the two-
// and one-launch iterations, which form the same values in different
// kernels, agree bit for bit.  Fused operations are written out (fma_t).
#pragma clang fp contract(off)

namespace {

constexpr int kBlock = 256;

enum { kRZ = 0, kPQ = 1, kAlpha = 2, kBeta = 3, kRR = 4, kBB = 5, kActive = 6, kIter = 7, kNS = 8 };

// Deterministic workgroup sums of N values at once: a butterfly within each
// 64-lane wavefront (every lane ends with the same bits: each step adds the
// same two operands, commuted), then the four wavefront sums in wave order.
// sh needs N * (kBlock / 64) doubles.
template <int N>
__device__ __forceinline__ void block_sum_n(double (&v)[N], double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int n = 0; n < N; ++n) v[n] += __shfl_xor(v[n], off, 64);
  constexpr int W = kBlock / 64;
  const int wv = threadIdx.x / 64;
  if (threadIdx.x % 64 == 0)
#pragma unroll
    for (int n = 0; n < N; ++n) sh[wv * N + n] = v[n];
  __syncthreads();
#pragma unroll
  for (int n = 0; n < N; ++n) {
    double t = sh[n];
#pragma unroll
    for (int w = 1; w < W; ++w) t += sh[w * N + n];
    v[n] = t;
  }
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ double block_sum(double v, double* sh) {
  double a[1] = {v};
  block_sum_n<1>(a, sh);
  return a[0];
}

// Rows [row0, row1) of A x, CSR-stream: the workgroup walks the rows'
// contiguous entry range [rp[row0], rp[row1]) in chunks of kChunk entries,
// each thread loading kPer entries of a chunk at once (coalesced values and
// column ids, then the x gathers), and stages the products in LDS; then the
// thread of row i (= row0 + threadIdx.x) adds its row's products in
// ascending entry order.  Returns that row sum (0 for threads past row1).
constexpr int kPer = 8;
constexpr int kChunk = kPer * kBlock;

// The loads of rows_spmv that do not depend on x, for the first chunk:
// issued early (before a kernel's scalar reductions) so that their latency
// overlaps; spmv_finish completes the product.
template <typename T>
struct SpmvHead {
  int32_t k0, k1, ra, rb;
  T a[kPer];
  int32_t c[kPer];
};

template <typename T>
__device__ __forceinline__ void spmv_begin(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                           const T* __restrict__ av, int32_t row0, int32_t row1, SpmvHead<T>& h) {
  const int32_t i = row0 + threadIdx.x;
  h.k0 = rp[row0];
  h.k1 = rp[row1];
  h.ra = i < row1 ? rp[i] : 0;
  h.rb = i < row1 ? rp[i + 1] : 0;
  const int32_t n = h.k1 - h.k0 < kChunk ? h.k1 - h.k0 : kChunk;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int32_t e = threadIdx.x + j * kBlock;
    h.a[j] = e < n ? av[h.k0 + e] : T(0);
    h.c[j] = e < n ? ci[h.k0 + e] : 0;
  }
}

template <typename T>
__device__ __forceinline__ T spmv_finish(const int32_t* __restrict__ ci, const T* __restrict__ av,
                                         const T* __restrict__ x, SpmvHead<T>& h, T* s_prod) {
  T acc = T(0);
  bool first = true;
  for (int32_t c0 = h.k0; c0 < h.k1; c0 += kChunk) {
    const int32_t n = h.k1 - c0 < kChunk ? h.k1 - c0 : kChunk;
    if (!first) {
#pragma unroll
      for (int j = 0; j < kPer; ++j) {
        const int32_t e = threadIdx.x + j * kBlock;
        h.a[j] = e < n ? av[c0 + e] : T(0);
        h.c[j] = e < n ? ci[c0 + e] : 0;
      }
    }
    first = false;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int32_t e = threadIdx.x + j * kBlock;
      if (e < n) s_prod[e] = h.a[j] * x[h.c[j]];
    }
    __syncthreads();
    const int32_t lo = h.ra > c0 ? h.ra : c0, hi = h.rb < c0 + n ? h.rb : c0 + n;
    for (int32_t e = lo; e < hi; ++e) acc += s_prod[e - c0];
    __syncthreads();
  }
  return acc;
}

template <typename T>
__device__ __forceinline__ T rows_spmv(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                       const T* __restrict__ av, const T* __restrict__ x, int32_t row0,
                                       int32_t row1, T* s_prod) {
  const int32_t i = row0 + threadIdx.x;
  const int32_t k0 = rp[row0], k1 = rp[row1];
  const int32_t ra = i < row1 ? rp[i] : 0, rb = i < row1 ? rp[i + 1] : 0;
  T acc = T(0);
  for (int32_t c0 = k0; c0 < k1; c0 += kChunk) {
    const int32_t n = k1 - c0 < kChunk ? k1 - c0 : kChunk;
    T a[kPer];
    int32_t c[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int32_t e = threadIdx.x + j * kBlock;
      a[j] = e < n ? av[c0 + e] : T(0);
      c[j] = e < n ? ci[c0 + e] : 0;
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int32_t e = threadIdx.x + j * kBlock;
      if (e < n) s_prod[e] = a[j] * x[c[j]];
    }
    __syncthreads();
    const int32_t lo = ra > c0 ? ra : c0, hi = rb < c0 + n ? rb : c0 + n;
    for (int32_t e = lo; e < hi; ++e) acc += s_prod[e - c0];
    __syncthreads();
  }
  return acc;
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
  __shared__ T s_prod[kChunk];
  const int32_t blk = (int32_t)tmh::xcd_block(blockIdx.x, gridDim.x);
  const int32_t row0 = blk_row0[blk], row1 = blk_row1[blk];
  const T ax = rows_spmv<T>(rp, ci, av, x, row0, row1, s_prod);
  const int32_t i = row0 + threadIdx.x;
  double rz = 0.0, bb = 0.0, rr = 0.0;
  if (i < row1) {
    const T ri = b[i] - ax;
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
  const int32_t blk = (int32_t)tmh::xcd_block(blockIdx.x, gridDim.x);
  const int32_t s = blk_sys[blk];
  if (sc[s * kNS + kActive] == 0.0) return;  // uniform per workgroup
  __shared__ T s_prod[kChunk];
  const int32_t row0 = blk_row0[blk], row1 = blk_row1[blk];
  const T qv = rows_spmv<T>(rp, ci, av, p, row0, row1, s_prod);
  const int32_t i = row0 + threadIdx.x;
  double pq = 0.0;
  if (i < row1) {
    const T qi = qv;
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
  const int32_t blk = (int32_t)tmh::xcd_block(blockIdx.x, gridDim.x);
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
  const int32_t blk = (int32_t)tmh::xcd_block(blockIdx.x, gridDim.x);
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

// a * b + c rounded once.  The two- and one-launch iterations write their
// vector updates with it so that no update depends on which products the
// compiler contracts: the one-launch step forms u = M r in registers where the
// two-launch one loads it, and a contraction of that product into p = u + beta p
// would round differently.
template <typename T>
__device__ __forceinline__ T fma_t(T a, T b, T c) {
  if constexpr (sizeof(T) == 8)
    return __builtin_fma(a, b, c);
  else
    return __builtin_fmaf(a, b, c);
}

// ---------------------------------------------------------------------------
// Two-launch PCG iteration (Chronopoulos-Gear recurrences, u = M r, w = A u,
// s = A p):
//   vec  (V_k): beta_k = gamma_k / gamma_{k-1},
//               alpha_k = gamma_k / (delta_k - beta_k gamma_k / alpha_{k-1})
//               (k = 0: beta = 0, alpha = gamma / delta);
//               p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s,
//               u = M r; partials of gamma_{k+1} = (r, u) and |r|^2
//   spmv (S_k): |r|^2 test (clears the system's active flag), w = A u,
//               partials of delta_{k+1} = (w, u)
// No kernel waits on another workgroup: every workgroup of the consuming
// kernel sums its system's partials itself, in workgroup order, so all of
// them derive bit-identical scalars without fences or atomics.  What one
// kernel both reads and writes is double-buffered by the iteration's parity
// P (V_k reads slot 1-P, writes slot P):
//   part[(0|1)*nb + w] gamma partials, part[2*nb + w] |r|^2, part[3*nb + w]
//   delta, part[4*nb + w] |b|^2;
//   sc[s*16 + 8 + 4*P + {0: alpha, 1: gamma, 2: iterations}].
// Host-visible copies (written by the system's first workgroup only):
//   sc[s*16 + {4: |r|^2, 5: |b|^2, 6: active, 7: iterations}].
// ---------------------------------------------------------------------------
constexpr int kNS2 = 16;


// sum over the system's workgroups [b0, b1) of part[w], in w order (every
// thread of the workgroup gets the same value)
__device__ __forceinline__ double system_sum(const double* __restrict__ part, int32_t b0, int32_t b1,
                                             double* sh) {
  double v = 0.0;
  for (int32_t w = b0 + threadIdx.x; w < b1; w += kBlock) v += part[w];
  return block_sum<double>(v, sh);
}

// The partials part[w], w in [b0, b1), that this thread adds in system_sum
// (w = b0 + threadIdx.x + k kBlock, ascending): the first kPartRegs of them
// loaded into registers up front (part_load), so that their latency overlaps
// the kernel's other loads, then added in the same order (part_add), which
// gives system_sum's bits after block_sum_n.
constexpr int kPartRegs = 8;  // 2048 workgroups per system in registers

struct PartRegs {
  double v[kPartRegs];
};

__device__ __forceinline__ void part_load(const double* __restrict__ part, int32_t b0, int32_t b1, PartRegs& r) {
#pragma unroll
  for (int k = 0; k < kPartRegs; ++k) {
    const int32_t w = b0 + (int32_t)threadIdx.x + k * kBlock;
    r.v[k] = w < b1 ? part[w] : 0.0;
  }
}

__device__ __forceinline__ double part_add(const double* __restrict__ part, int32_t b0, int32_t b1,
                                           const PartRegs& r) {
  double v = 0.0;
#pragma unroll
  for (int k = 0; k < kPartRegs; ++k)
    if (b0 + (int32_t)threadIdx.x + k * kBlock < b1) v += r.v[k];
  for (int32_t w = b0 + (int32_t)threadIdx.x + kPartRegs * kBlock; w < b1; w += kBlock) v += part[w];
  return v;
}

// two system sums in one workgroup reduction
__device__ __forceinline__ void system_sum2(const double* __restrict__ pa, const double* __restrict__ pb, int32_t b0,
                                            int32_t b1, double* sh, double& a, double& b) {
  double v[2] = {0.0, 0.0};
  for (int32_t w = b0 + threadIdx.x; w < b1; w += kBlock) {
    v[0] += pa[w];
    v[1] += pb[w];
  }
  block_sum_n<2>(v, sh);
  a = v[0];
  b = v[1];
}

// r = b - A x, u = dinv r, p = s = 0; partials gamma_0 (slot 1), |r|^2, |b|^2
template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg2_init_kernel(const int32_t* __restrict__ blk_sys, const int32_t* __restrict__ blk_row0,
                    const int32_t* __restrict__ blk_row1, const int32_t* __restrict__ sys_blk,
                    const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const T* __restrict__ av,
                    const T* __restrict__ b, const T* __restrict__ x, const T* __restrict__ dinv, T* __restrict__ r,
                    T* __restrict__ u, T* __restrict__ p, T* __restrict__ sv, double* __restrict__ part,
                    double* __restrict__ sc, int32_t nb) {
  __shared__ double sh[kBlock];
  __shared__ T s_prod[kChunk];
  const int32_t blk = (int32_t)tmh::xcd_block(blockIdx.x, gridDim.x);
  const int32_t sy = blk_sys[blk];
  const int32_t row0 = blk_row0[blk], row1 = blk_row1[blk];
  const T ax = rows_spmv<T>(rp, ci, av, x, row0, row1, s_prod);
  const int32_t i = row0 + threadIdx.x;
  double g = 0.0, bb = 0.0, rr = 0.0;
  if (i < row1) {
    const T ri = b[i] - ax;
    const T ui = dinv[i] * ri;
    r[i] = ri;
    u[i] = ui;
    p[i] = T(0);
    sv[i] = T(0);
    g = (double)ri * (double)ui;
    bb = (double)b[i] * (double)b[i];
    rr = (double)ri * (double)ri;
  }
  {
    double v[3] = {g, bb, rr};
    block_sum_n<3>(v, sh);
    g = v[0];
    bb = v[1];
    rr = v[2];
  }
  if (threadIdx.x == 0) {
    part[1 * nb + blk] = g;
    part[2 * nb + blk] = rr;
    part[4 * nb + blk] = bb;
    if (blk == sys_blk[sy]) {
      sc[sy * kNS2 + 8 + 4 + 2] = 0.0;  // iterations in slot 1 (read by V_0)
      sc[sy * kNS2 + 7] = 0.0;
    }
  }
}

// |r|^2 test, then w = A u and partials of delta = (w, u).  first: also sums
// |b|^2 (published for the later S launches, which read it from sc).
template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg2_spmv_kernel(const int32_t* __restrict__ blk_sys, const int32_t* __restrict__ blk_row0,
                    const int32_t* __restrict__ blk_row1, const int32_t* __restrict__ sys_blk,
                    const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const T* __restrict__ av,
                    const T* __restrict__ u, T* __restrict__ w, double* __restrict__ part, double* __restrict__ sc,
                    int32_t nb, double tol2, int first) {
  __shared__ double sh[kBlock];
  const int32_t blk = (int32_t)tmh::xcd_block(blockIdx.x, gridDim.x);
  const int32_t sy = blk_sys[blk];
  const int32_t b0 = sys_blk[sy], b1 = sys_blk[sy + 1];
  double* c = sc + sy * kNS2;
  // the partials, the row's u and the matrix loads all go out before the
  // |r|^2 reduction
  const int32_t row0 = blk_row0[blk], row1 = blk_row1[blk];
  const int32_t i = row0 + threadIdx.x;
  PartRegs pr, pb;
  part_load(part + 2 * nb, b0, b1, pr);
  if (first) part_load(part + 4 * nb, b0, b1, pb);
  const T ui = i < row1 ? u[i] : T(0);
  SpmvHead<T> head;
  spmv_begin<T>(rp, ci, av, row0, row1, head);
  double rr, bb;
  if (first) {
    double v[2] = {part_add(part + 2 * nb, b0, b1, pr), part_add(part + 4 * nb, b0, b1, pb)};
    block_sum_n<2>(v, sh);
    rr = v[0];
    bb = v[1];
  } else {
    double v[1] = {part_add(part + 2 * nb, b0, b1, pr)};
    block_sum_n<1>(v, sh);
    rr = v[0];
    bb = c[5];
  }
  const bool active = rr > tol2 * bb && bb > 0.0;
  if (threadIdx.x == 0 && blk == b0) {
    c[4] = rr;
    if (first) c[5] = bb;
    c[6] = active ? 1.0 : 0.0;
  }
  if (!active) return;  // uniform per workgroup
  __shared__ T s_prod[kChunk];
  const T wv = spmv_finish<T>(ci, av, u, head, s_prod);
  double d = 0.0;
  if (i < row1) {
    w[i] = wv;
    d = (double)wv * (double)ui;
  }
  d = block_sum<T>(d, sh);
  if (threadIdx.x == 0) part[3 * nb + blk] = d;
}

// vector updates of iteration k (parity P); partials of gamma_{k+1}, |r|^2
template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg2_vec_kernel(const int32_t* __restrict__ blk_sys, const int32_t* __restrict__ blk_row0,
                   const int32_t* __restrict__ blk_row1, const int32_t* __restrict__ sys_blk,
                   const T* __restrict__ dinv, const T* __restrict__ w, T* __restrict__ x, T* __restrict__ r,
                   T* __restrict__ u, T* __restrict__ p, T* __restrict__ sv, double* __restrict__ part,
                   double* __restrict__ sc, int32_t nb, int parity) {
  __shared__ double sh[kBlock];
  const int32_t blk = (int32_t)tmh::xcd_block(blockIdx.x, gridDim.x);
  const int32_t sy = blk_sys[blk];
  double* c = sc + sy * kNS2;
  const int32_t b0 = sys_blk[sy], b1 = sys_blk[sy + 1];
  const int P = parity & 1, Q = P ^ 1;
  // the row's vector entries go out with the active flag and the partials
  const int32_t i = blk_row0[blk] + threadIdx.x;
  const bool has = i < blk_row1[blk];
  T u0 = T(0), p0 = T(0), w0 = T(0), s0 = T(0), x0 = T(0), r0 = T(0), d0 = T(0);
  if (has) {
    u0 = u[i];
    p0 = p[i];
    w0 = w[i];
    s0 = sv[i];
    x0 = x[i];
    r0 = r[i];
    d0 = dinv[i];
  }
  PartRegs pg, pd;
  part_load(part + Q * nb, b0, b1, pg);
  part_load(part + 3 * nb, b0, b1, pd);
  if (c[6] == 0.0) return;  // cleared by the preceding S launch; uniform per system
  double gd[2] = {part_add(part + Q * nb, b0, b1, pg), part_add(part + 3 * nb, b0, b1, pd)};
  block_sum_n<2>(gd, sh);
  const double gamma = gd[0], delta = gd[1];
  const double* prev = c + 8 + 4 * Q;
  const double iter = prev[2];
  double alpha, beta;
  if (iter == 0.0) {
    beta = 0.0;
    alpha = delta != 0.0 ? gamma / delta : 0.0;
  } else {
    beta = prev[1] != 0.0 ? gamma / prev[1] : 0.0;
    const double den = delta - beta * gamma / prev[0];
    alpha = den != 0.0 ? gamma / den : 0.0;
  }
  if (threadIdx.x == 0 && blk == b0) {
    double* cur = c + 8 + 4 * P;
    cur[0] = alpha;
    cur[1] = gamma;
    cur[2] = iter + 1.0;
    c[7] = iter + 1.0;
  }
  const T al = (T)alpha, be = (T)beta;
  double g = 0.0, rr = 0.0;
  if (has) {
    const T pi = fma_t(be, p0, u0);
    const T si = fma_t(be, s0, w0);
    p[i] = pi;
    sv[i] = si;
    x[i] = fma_t(al, pi, x0);
    const T ri = fma_t(-al, si, r0);
    const T ui = d0 * ri;
    r[i] = ri;
    u[i] = ui;
    g = (double)ri * (double)ui;
    rr = (double)ri * (double)ri;
  }
  {
    double v[2] = {g, rr};
    block_sum_n<2>(v, sh);
    g = v[0];
    rr = v[1];
  }
  if (threadIdx.x == 0) {
    part[P * nb + blk] = g;
    part[2 * nb + blk] = rr;
  }
}

// ---------------------------------------------------------------------------
// One-launch PCG iteration: the recurrences and the arithmetic of the
// two-launch one, with the vector update folded into the SpMV launch.  Step k
// (parity P, reading slot Q = P ^ 1 of the double-buffered r, A u, A p):
//   scalars: |r_Q|^2 test, gamma = (r_Q, u_Q), delta = (A u_Q, u_Q) from the
//            workgroup partials of step k-1, alpha / beta as the vec kernel;
//   own rows: p = u_Q + beta p, s_P = w_Q + beta s_Q, x += alpha p,
//            r_P = r_Q - alpha s_P, u_P = M r_P;
//   w_P = A u_P, where the u_P of every column is formed from slot Q by the
//            same inline arithmetic (bit-identical to the owner's);
//   partials of gamma, delta and |r_P|^2 for step k+1.
// Partials: part[(0|1) nb] gamma, [(2|3) nb] delta, [(4|5) nb] |r|^2, [6 nb]
// |b|^2, by slot; scalars as the two-launch variant (the active flag, once
// cleared, stops every later step).  Every iterate is bit-identical to the
// two-launch one.
// ---------------------------------------------------------------------------

// next u = M (r - alpha (w + beta s)) of one row, with s and r written back
template <typename T>
__device__ __forceinline__ T cg1_next(T r, T w, T s, T d, T al, T be, T& s_new, T& r_new) {
  s_new = fma_t(be, s, w);
  r_new = fma_t(-al, s_new, r);
  return d * r_new;
}

// r = b - A x, s = 0, p = 0 (slot 0); partials gamma_0, |r|^2, |b|^2
template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg1_init_kernel(const int32_t* __restrict__ blk_sys, const int32_t* __restrict__ blk_row0,
                    const int32_t* __restrict__ blk_row1, const int32_t* __restrict__ sys_blk,
                    const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const T* __restrict__ av,
                    const T* __restrict__ b, const T* __restrict__ x, const T* __restrict__ dinv, T* __restrict__ r,
                    T* __restrict__ p, T* __restrict__ sv, double* __restrict__ part, double* __restrict__ sc,
                    int32_t nb) {
  __shared__ double sh[3 * (kBlock / 64)];
  __shared__ T s_prod[kChunk];
  const int32_t blk = (int32_t)tmh::xcd_block(blockIdx.x, gridDim.x);
  const int32_t sy = blk_sys[blk];
  const int32_t row0 = blk_row0[blk], row1 = blk_row1[blk];
  const T ax = rows_spmv<T>(rp, ci, av, x, row0, row1, s_prod);
  const int32_t i = row0 + threadIdx.x;
  double g = 0.0, bb = 0.0, rr = 0.0;
  if (i < row1) {
    const T ri = b[i] - ax;
    const T ui = dinv[i] * ri;
    r[i] = ri;
    p[i] = T(0);
    sv[i] = T(0);
    g = (double)ri * (double)ui;
    bb = (double)b[i] * (double)b[i];
    rr = (double)ri * (double)ri;
  }
  double v[3] = {g, bb, rr};
  block_sum_n<3>(v, sh);
  if (threadIdx.x == 0) {
    part[0 * nb + blk] = v[0];
    part[4 * nb + blk] = v[2];
    part[6 * nb + blk] = v[1];
    if (blk == sys_blk[sy]) {
      sc[sy * kNS2 + 8 + 0 + 2] = 0.0;  // iterations in slot 0 (read by step 0)
      sc[sy * kNS2 + 7] = 0.0;
    }
  }
}

// w = A u of slot 0 (u = M r formed per column), partials delta_0; publishes
// |b|^2, |r|^2 and the active flag
template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg1_spmv0_kernel(const int32_t* __restrict__ blk_sys, const int32_t* __restrict__ blk_row0,
                     const int32_t* __restrict__ blk_row1, const int32_t* __restrict__ sys_blk,
                     const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const T* __restrict__ av,
                     const T* __restrict__ dinv, const T* __restrict__ r, T* __restrict__ w,
                     double* __restrict__ part, double* __restrict__ sc, int32_t nb, double tol2) {
  __shared__ double sh[2 * (kBlock / 64)];
  __shared__ T s_prod[kChunk];
  const int32_t blk = (int32_t)tmh::xcd_block(blockIdx.x, gridDim.x);
  const int32_t sy = blk_sys[blk];
  const int32_t b0 = sys_blk[sy], b1 = sys_blk[sy + 1];
  double* c = sc + sy * kNS2;
  double rr, bb;
  system_sum2(part + 4 * nb, part + 6 * nb, b0, b1, sh, rr, bb);
  const bool active = rr > tol2 * bb && bb > 0.0;
  if (threadIdx.x == 0 && blk == b0) {
    c[4] = rr;
    c[5] = bb;
    c[6] = active ? 1.0 : 0.0;
  }
  if (!active) return;  // uniform per workgroup
  const int32_t row0 = blk_row0[blk], row1 = blk_row1[blk];
  const int32_t i = row0 + threadIdx.x;
  const int32_t k0 = rp[row0], k1 = rp[row1];
  const int32_t ra = i < row1 ? rp[i] : 0, rb = i < row1 ? rp[i + 1] : 0;
  T acc = T(0);
  for (int32_t c0 = k0; c0 < k1; c0 += kChunk) {
    const int32_t n = k1 - c0 < kChunk ? k1 - c0 : kChunk;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int32_t e = threadIdx.x + j * kBlock;
      if (e < n) {
        const int32_t col = ci[c0 + e];
        s_prod[e] = av[c0 + e] * (dinv[col] * r[col]);
      }
    }
    __syncthreads();
    const int32_t lo = ra > c0 ? ra : c0, hi = rb < c0 + n ? rb : c0 + n;
    for (int32_t e = lo; e < hi; ++e) acc += s_prod[e - c0];
    __syncthreads();
  }
  double d = 0.0;
  if (i < row1) {
    w[i] = acc;
    d = (double)acc * (double)(dinv[i] * r[i]);
  }
  d = block_sum<T>(d, sh);
  if (threadIdx.x == 0) part[2 * nb + blk] = d;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg1_step_kernel(const int32_t* __restrict__ blk_sys, const int32_t* __restrict__ blk_row0,
                    const int32_t* __restrict__ blk_row1, const int32_t* __restrict__ sys_blk,
                    const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const T* __restrict__ av,
                    const T* __restrict__ dinv, const T* __restrict__ rQ, const T* __restrict__ wQ,
                    const T* __restrict__ sQ, T* __restrict__ rP, T* __restrict__ wP, T* __restrict__ sP,
                    T* __restrict__ p, T* __restrict__ x, double* __restrict__ part, double* __restrict__ sc,
                    int32_t nb, double tol2, int parity) {
  __shared__ double sh[3 * (kBlock / 64)];
  __shared__ T s_prod[kChunk];
  __shared__ int s_run;
  const int32_t blk = (int32_t)tmh::xcd_block(blockIdx.x, gridDim.x);
  const int32_t sy = blk_sys[blk];
  const int32_t b0 = sys_blk[sy], b1 = sys_blk[sy + 1];
  double* c = sc + sy * kNS2;
  const int P = parity & 1, Q = P ^ 1;
  // a system that stopped stays stopped: once it is inactive its slots are no
  // longer written, so the parity would otherwise bring back an older |r|^2.
  // One thread reads the flag (the system's first workgroup may be clearing it
  // in this very step) and the workgroup follows that one read.
  if (threadIdx.x == 0) s_run = c[6] != 0.0;
  const int32_t row0 = blk_row0[blk], row1 = blk_row1[blk];
  const int32_t i = row0 + threadIdx.x;
  const bool has = i < row1;
  // own-row state and the first SpMV chunk go out before the scalar reductions
  T rq = T(0), wq = T(0), sq = T(0), pq = T(0), xq = T(0), dq = T(0);
  if (has) {
    rq = rQ[i];
    wq = wQ[i];
    sq = sQ[i];
    pq = p[i];
    xq = x[i];
    dq = dinv[i];
  }
  const int32_t k0 = rp[row0], k1 = rp[row1];
  const int32_t ra = has ? rp[i] : 0, rb = has ? rp[i + 1] : 0;
  T a[kPer], gr[kPer], gw[kPer], gs[kPer], gd[kPer];
  {
    const int32_t n = k1 - k0 < kChunk ? k1 - k0 : kChunk;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int32_t e = threadIdx.x + j * kBlock;
      a[j] = T(0);
      gr[j] = gw[j] = gs[j] = gd[j] = T(0);
      if (e < n) {
        const int32_t col = ci[k0 + e];
        a[j] = av[k0 + e];
        gr[j] = rQ[col];
        gw[j] = wQ[col];
        gs[j] = sQ[col];
        gd[j] = dinv[col];
      }
    }
  }
  __syncthreads();
  if (!s_run) return;  // uniform per workgroup
  double sums[3] = {0.0, 0.0, 0.0};
  {
    // gamma_Q, delta_Q, |r_Q|^2 over the system's workgroups, in workgroup order
    for (int32_t wgi = b0 + threadIdx.x; wgi < b1; wgi += kBlock) {
      sums[0] += part[Q * nb + wgi];
      sums[1] += part[(2 + Q) * nb + wgi];
      sums[2] += part[(4 + Q) * nb + wgi];
    }
    block_sum_n<3>(sums, sh);
  }
  const double gamma = sums[0], delta = sums[1], rr = sums[2];
  const double bb = c[5];
  const bool active = rr > tol2 * bb && bb > 0.0;
  if (threadIdx.x == 0 && blk == b0) {
    c[4] = rr;
    c[6] = active ? 1.0 : 0.0;
  }
  if (!active) return;  // uniform per workgroup
  const double* prev = c + 8 + 4 * Q;
  const double iter = prev[2];
  double alpha, beta;
  if (iter == 0.0) {
    beta = 0.0;
    alpha = delta != 0.0 ? gamma / delta : 0.0;
  } else {
    beta = prev[1] != 0.0 ? gamma / prev[1] : 0.0;
    const double den = delta - beta * gamma / prev[0];
    alpha = den != 0.0 ? gamma / den : 0.0;
  }
  if (threadIdx.x == 0 && blk == b0) {
    double* cur = c + 8 + 4 * P;
    cur[0] = alpha;
    cur[1] = gamma;
    cur[2] = iter + 1.0;
    c[7] = iter + 1.0;
  }
  const T al = (T)alpha, be = (T)beta;
  double g = 0.0, r2 = 0.0;
  T ui = T(0);
  if (has) {
    const T uq = dq * rq;
    const T pi = fma_t(be, pq, uq);
    T si, ri;
    ui = cg1_next<T>(rq, wq, sq, dq, al, be, si, ri);
    p[i] = pi;
    sP[i] = si;
    x[i] = fma_t(al, pi, xq);
    rP[i] = ri;
    g = (double)ri * (double)ui;
    r2 = (double)ri * (double)ri;
  }
  // w_P = A u_P, rows summed in ascending entry order (CSR-stream through LDS)
  T acc = T(0);
  bool first = true;
  for (int32_t c0 = k0; c0 < k1; c0 += kChunk) {
    const int32_t n = k1 - c0 < kChunk ? k1 - c0 : kChunk;
    if (!first) {
#pragma unroll
      for (int j = 0; j < kPer; ++j) {
        const int32_t e = threadIdx.x + j * kBlock;
        if (e < n) {
          const int32_t col = ci[c0 + e];
          a[j] = av[c0 + e];
          gr[j] = rQ[col];
          gw[j] = wQ[col];
          gs[j] = sQ[col];
          gd[j] = dinv[col];
        }
      }
    }
    first = false;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int32_t e = threadIdx.x + j * kBlock;
      if (e < n) {
        T sn, rn;
        const T uc = cg1_next<T>(gr[j], gw[j], gs[j], gd[j], al, be, sn, rn);
        s_prod[e] = a[j] * uc;
      }
    }
    __syncthreads();
    const int32_t lo = ra > c0 ? ra : c0, hi = rb < c0 + n ? rb : c0 + n;
    for (int32_t e = lo; e < hi; ++e) acc += s_prod[e - c0];
    __syncthreads();
  }
  double d = 0.0;
  if (has) {
    wP[i] = acc;
    d = (double)acc * (double)ui;
  }
  double v[3] = {g, d, r2};
  block_sum_n<3>(v, sh);
  if (threadIdx.x == 0) {
    part[P * nb + blk] = v[0];
    part[(2 + P) * nb + blk] = v[1];
    part[(4 + P) * nb + blk] = v[2];
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    spmv_kernel(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const T* __restrict__ av,
                const T* __restrict__ x, T* __restrict__ y, int64_t n_rows) {
  // 256 rows per workgroup, CSR-stream
  __shared__ T s_prod[kChunk];
  const int64_t r0 = tmh::xcd_block(blockIdx.x, gridDim.x) * kBlock;
  const int32_t row0 = (int32_t)r0, row1 = (int32_t)(r0 + kBlock < n_rows ? r0 + kBlock : n_rows);
  const T v = rows_spmv<T>(rp, ci, av, x, row0, row1, s_prod);
  const int32_t i = row0 + threadIdx.x;
  if (i < row1) y[i] = v;
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

template <typename T>
int cg2_init(const tm_pcg_view* v, const void* b, const void* x, hipStream_t s) {
  const int32_t nb = (int32_t)v->n_blocks;
  const double tol2 = v->rtol * v->rtol;
  cg2_init_kernel<T><<<nb, kBlock, 0, s>>>(v->blk_sys, v->blk_row0, v->blk_row1, v->sys_blk, v->row_ptr, v->col_idx,
                                           (const T*)v->values, (const T*)b, (const T*)x, (const T*)v->dinv,
                                           (T*)v->r, (T*)v->z, (T*)v->p, (T*)v->s, v->partials, v->scalars, nb);
  cg2_spmv_kernel<T><<<nb, kBlock, 0, s>>>(v->blk_sys, v->blk_row0, v->blk_row1, v->sys_blk, v->row_ptr, v->col_idx,
                                           (const T*)v->values, (const T*)v->z, (T*)v->q, v->partials, v->scalars,
                                           nb, tol2, 1);
  return hip_status("pcg init");
}

template <typename T>
int cg2_iterate(const tm_pcg_view* v, void* x, int32_t n_iter, hipStream_t s) {
  const int32_t nb = (int32_t)v->n_blocks;
  const double tol2 = v->rtol * v->rtol;
  for (int32_t it = 0; it < n_iter; ++it) {
    cg2_vec_kernel<T><<<nb, kBlock, 0, s>>>(v->blk_sys, v->blk_row0, v->blk_row1, v->sys_blk, (const T*)v->dinv,
                                            (const T*)v->q, (T*)x, (T*)v->r, (T*)v->z, (T*)v->p, (T*)v->s,
                                            v->partials, v->scalars, nb, (v->parity + it) & 1);
    cg2_spmv_kernel<T><<<nb, kBlock, 0, s>>>(v->blk_sys, v->blk_row0, v->blk_row1, v->sys_blk, v->row_ptr,
                                             v->col_idx, (const T*)v->values, (const T*)v->z, (T*)v->q, v->partials,
                                             v->scalars, nb, tol2, 0);
  }
  return hip_status("pcg iterate");
}

template <typename T>
int cg1_init(const tm_pcg_view* v, const void* b, const void* x, hipStream_t s) {
  const int32_t nb = (int32_t)v->n_blocks;
  const double tol2 = v->rtol * v->rtol;
  cg1_init_kernel<T><<<nb, kBlock, 0, s>>>(v->blk_sys, v->blk_row0, v->blk_row1, v->sys_blk, v->row_ptr, v->col_idx,
                                           (const T*)v->values, (const T*)b, (const T*)x, (const T*)v->dinv,
                                           (T*)v->r, (T*)v->p, (T*)v->s, v->partials, v->scalars, nb);
  cg1_spmv0_kernel<T><<<nb, kBlock, 0, s>>>(v->blk_sys, v->blk_row0, v->blk_row1, v->sys_blk, v->row_ptr,
                                            v->col_idx, (const T*)v->values, (const T*)v->dinv, (const T*)v->r,
                                            (T*)v->q, v->partials, v->scalars, nb, tol2);
  return hip_status("pcg init");
}

// slot 0: r, q (A u), s (A p); slot 1: z, aux[0, n), aux[n, 2n)
template <typename T>
int cg1_iterate(const tm_pcg_view* v, void* x, int32_t n_iter, hipStream_t s) {
  const int32_t nb = (int32_t)v->n_blocks;
  const double tol2 = v->rtol * v->rtol;
  T* rb[2] = {(T*)v->r, (T*)v->z};
  T* wb[2] = {(T*)v->q, (T*)v->aux};
  T* sb[2] = {(T*)v->s, (T*)v->aux + v->n_rows};
  for (int32_t it = 0; it < n_iter; ++it) {
    const int Q = (v->parity + it) & 1, P = Q ^ 1;
    cg1_step_kernel<T><<<nb, kBlock, 0, s>>>(v->blk_sys, v->blk_row0, v->blk_row1, v->sys_blk, v->row_ptr,
                                             v->col_idx, (const T*)v->values, (const T*)v->dinv, rb[Q], wb[Q],
                                             sb[Q], rb[P], wb[P], sb[P], (T*)v->p, (T*)x, v->partials, v->scalars,
                                             nb, tol2, P);
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
  if (v->two_launch < 0 || v->two_launch > 2) return tmh::fail(TM_ERR_ARG, "unknown PCG variant");
  if (v->two_launch >= 1 && !v->s) return tmh::fail(TM_ERR_ARG, "the two- and one-launch PCG need s");
  if (v->two_launch == 2 && !v->aux) return tmh::fail(TM_ERR_ARG, "the one-launch PCG needs aux [2 n_rows]");
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
  const int var = view->two_launch;
  if (view->dtype == TM_F64)
    return var == 2 ? cg1_init<double>(view, b, x, s) : var ? cg2_init<double>(view, b, x, s)
                                                           : pcg_init<double>(view, b, x, s);
  if (view->dtype == TM_F32)
    return var == 2 ? cg1_init<float>(view, b, x, s) : var ? cg2_init<float>(view, b, x, s)
                                                          : pcg_init<float>(view, b, x, s);
  return tmh::fail(TM_ERR_ARG, "unknown dtype");
  TM_TRY_END
}

extern "C" int tm_pcg_iterate(const tm_pcg_view* view, void* x, int32_t n_iter, tm_stream_t stream_) {
  TM_TRY_BEGIN
  int st = check_view(view);
  if (st) return st;
  if (!x || n_iter < 0) return tmh::fail(TM_ERR_ARG, "bad x / n_iter");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream_);
  const int var = view->two_launch;
  if (view->dtype == TM_F64)
    return var == 2 ? cg1_iterate<double>(view, x, n_iter, s) : var ? cg2_iterate<double>(view, x, n_iter, s)
                                                                   : pcg_iterate<double>(view, x, n_iter, s);
  if (view->dtype == TM_F32)
    return var == 2 ? cg1_iterate<float>(view, x, n_iter, s) : var ? cg2_iterate<float>(view, x, n_iter, s)
                                                                  : pcg_iterate<float>(view, x, n_iter, s);
  return tmh::fail(TM_ERR_ARG, "unknown dtype");
  TM_TRY_END
}
