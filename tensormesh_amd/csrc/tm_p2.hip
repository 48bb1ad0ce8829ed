// Stage G for quadratic simplices and block (vector) forms: one wavefront per
// element.
//
// Replaces the same reference chain as tm_p1.hip -- eval_shape_grad
// (tensormesh/element/element.py:956-968), jxw (transformation.py:1265-1272),
// the vmap'd forward and _integrate (element_assembler.py:190-209,416-438) --
// for triangle6 / tetra10 and for LinearElasticityElementAssembler
// (builtin.py:172-178), where the Jacobian varies per quadrature point and a
// local matrix has up to 30 x 30 entries.
//
// Work split inside the wave:
//   1. lanes load the element's B*D coordinates into LDS;
//   2. lane q builds J_q[i][j] = sum_b x[b][j] dphi_q[i][b], its inverse and
//      jxw_q = w_q |det J_q|;
//   3. lanes over (q, b, d) form grad_q[b][d] = sum_i dphi_q[i][b] Jinv_q[d][i];
//   4. lanes stride over the B*B*H*H local entries; entry (i, j, a, c) is
//      sum_q jxw_q * f_q(i, j, a, c) and is written to the slot buffer at
//      ((e*B + i)*B + j)*H*H + a*H + c, the flattening of the reference's
//      [E, B, B, H, H] element tensor.
// Linear elasticity uses Ba^T C Bb in closed form:
//   (Ba^T C Bb)[a][c] = lam g_i[a] g_j[c] + mu g_i[c] g_j[a] + mu delta_ac g_i.g_j
// which is the product of the Voigt matrices of functional/elasticity.py:132-284.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "tm_device.h"
#include "tm_internal.h"

namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kMaxQ = 16;      // Neo-Hookean kernel
constexpr int kMaxQWave = 32;  // element-matrix kernel (a 3 x 3 x 3 hexahedron rule has 27 points)


// MQ: largest number of quadrature points the LDS arrays hold.  upper: only
// the blocks i <= j (symmetric forms, TM_FLAG_UPPER); products are formed so
// that block (j, i) would be the exact transpose of block (i, j).
template <typename T, int D, int B, int H, int FORM, int MQ, int EPW>
__global__ void __launch_bounds__(kWave * kWavesPerBlock)
    wave_element_kernel(const int32_t* __restrict__ conn, const T* __restrict__ pts,
                        const T* __restrict__ coeff, const double* __restrict__ tab, int Q,
                        double lam, double mu, int64_t n_elem, int upper, T* __restrict__ slot,
                        const int32_t* __restrict__ slot_dest) {
  constexpr int L = kWave / EPW;                 // lanes per element
  constexpr int kEpb = kWavesPerBlock * EPW;     // elements per workgroup
  __shared__ double s_tab[MQ * (1 + B + D * B)];
  __shared__ uint16_t s_pair[B * B];
  __shared__ double s_x[kEpb][B * D];
  __shared__ double s_inv[kEpb][MQ][D * D];
  __shared__ double s_jxw[kEpb][MQ];
  __shared__ double s_kq[kEpb][MQ];
  __shared__ double s_g[kEpb][MQ * B * D];
  __shared__ int s_v[kEpb][B];

  static_assert(B <= L && MQ <= L, "an element's lanes index its nodes and its quadrature points");
  const int lane = threadIdx.x % L;   // lane within the element's group
  const int wv = threadIdx.x / L;     // element slot within the workgroup
  const int64_t e = (int64_t)blockIdx.x * kEpb + wv;
  const bool active = e < n_elem;

  const int tab_len = Q + Q * B + Q * D * B;
  for (int t = threadIdx.x; t < tab_len; t += blockDim.x) s_tab[t] = tab[t];
  const int n_pairs = tmh::fill_pairs<B>(s_pair, upper != 0);
  const double* w = s_tab;
  const double* phi = s_tab + Q;
  const double* dphi = s_tab + Q + Q * B;

  if (active) {
    if (lane < B) s_v[wv][lane] = conn[e * B + lane];
  }
  // the lane's block destinations (contribution order), loaded now so that
  // their latency hides behind the per-point stages
  constexpr int kPairRounds = (B * B + L - 1) / L;
  int32_t dst[kPairRounds];
#pragma unroll
  for (int r = 0; r < kPairRounds; ++r) {
    const int p = lane + r * L;
    dst[r] = (FORM == TM_FORM_LINEAR_ELASTICITY && slot_dest && active && p < n_pairs)
                 ? slot_dest[e * n_pairs + p] : 0;
  }
  __syncthreads();
  if (active) {
    for (int t = lane; t < B * D; t += L) {
      const int b = t / D, d = t % D;
      s_x[wv][t] = (double)pts[(int64_t)s_v[wv][b] * D + d];
    }
  }
  __syncthreads();
  // the Jacobians J[i][j] = sum_b dphi_b[i] x_b[j] (into s_inv, which
  // point q's lane then overwrites with the inverse)
  if (active)
    tmh::point_sums<D, B, L, MQ>(
        lane, Q, false, [&](int q, int i, int b) { return dphi[(q * D + i) * B + b]; },
        [&](int, int b, int j) { return s_x[wv][b * D + j]; },
        [&](int q, int i, int j) -> double& { return s_inv[wv][q][i * D + j]; });
  __syncthreads();
  if (active && lane < Q) {
    const int q = lane;
    double J[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) J[i][j] = s_inv[wv][q][i * D + j];
    double inv[D][D];
    double det;
    if constexpr (D == 2) {
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      inv[0][0] = J[1][1] / det;
      inv[0][1] = -J[0][1] / det;
      inv[1][0] = -J[1][0] / det;
      inv[1][1] = J[0][0] / det;
    } else {
      const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
      const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      const double r = 1.0 / det;
      inv[0][0] = c00 * r;
      inv[1][0] = c01 * r;
      inv[2][0] = c02 * r;
      inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r;
      inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r;
      inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r;
      inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
      inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r;
      inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
    }
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) s_inv[wv][q][i * D + j] = inv[i][j];
    s_jxw[wv][q] = w[q] * fabs(det);
    double kq = 1.0;
    if constexpr (FORM == TM_FORM_DIFFUSION_ELEM) kq = (double)coeff[e];
    if constexpr (FORM == TM_FORM_DIFFUSION_NODE) {
      kq = 0.0;
      for (int b = 0; b < B; ++b) kq += (double)coeff[s_v[wv][b]] * phi[q * B + b];
    }
    s_kq[wv][q] = kq;
  }
  __syncthreads();
  if (active) {
    for (int t = lane; t < Q * B * D; t += L) {
      const int q = t / (B * D), b = (t / D) % B, d = t % D;
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < D; ++i) s += dphi[(q * D + i) * B + b] * s_inv[wv][q][d * D + i];
      s_g[wv][t] = s;
    }
  }
  __syncthreads();
  if (!active) return;
  constexpr int HH = H * H;
  // pair p of the table is block (i, j): full blocks at p = i*B + j, upper
  // blocks packed at their index among the pairs i <= j
  T* out = slot + e * (int64_t)(n_pairs * HH);
  if constexpr (FORM == TM_FORM_LINEAR_ELASTICITY) {
    // one lane per node pair: the H x H block from one read of g_i, g_j
#pragma unroll
    for (int r = 0; r < kPairRounds; ++r) {
      const int p = lane + r * L;
      if (p >= n_pairs) break;
      int i = s_pair[p] >> 8, j = s_pair[p] & 255;
      // contribution-ordered slot, transposed where the pair is read as
      // (j, i): formed as block (j, i), the exact transpose of (i, j)
      T* o = out + p * HH;
      if (slot_dest) {
        const int32_t d = dst[r];
        o = slot + (int64_t)(d & INT32_MAX) * HH;
        if (d < 0) {
          const int t = i;
          i = j;
          j = t;
        }
      }
      double acc[HH];
#pragma unroll
      for (int k = 0; k < HH; ++k) acc[k] = 0.0;
      for (int q = 0; q < Q; ++q) {
        const double* g = &s_g[wv][q * B * D];
        double gi[D], gj[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
          gi[d] = g[i * D + d];
          gj[d] = g[j * D + d];
        }
        double dot = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) dot += gi[d] * gj[d];
        const double w = s_jxw[wv][q];
#pragma unroll
        for (int a = 0; a < H; ++a)
#pragma unroll
          for (int c = 0; c < H; ++c) {
            double t = lam * (gi[a] * gj[c]) + mu * (gi[c] * gj[a]);
            if (a == c) t += mu * dot;
            acc[a * H + c] += t * w;
          }
      }
#pragma unroll
      for (int k = 0; k < HH; ++k) o[k] = (T)acc[k];
    }
    return;
  }
  static_assert(FORM == TM_FORM_LINEAR_ELASTICITY || H == 1, "scalar forms have one value per pair");
  for (int p = lane; p < n_pairs; p += L) {
    const int i = s_pair[p] >> 8, j = s_pair[p] & 255;
    double acc = 0.0;
    for (int q = 0; q < Q; ++q) {
      const double* g = &s_g[wv][q * B * D];
      double val;
      if constexpr (FORM == TM_FORM_MASS) {
        val = phi[q * B + i] * phi[q * B + j];
      } else {
        double dot = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) dot += g[i * D + d] * g[j * D + d];
        val = s_kq[wv][q] * dot;
      }
      acc += val * s_jxw[wv][q];
    }
    if (slot_dest)
      slot[slot_dest[e * n_pairs + p] & INT32_MAX] = (T)acc;
    else
      out[p] = (T)acc;
  }
}

// Neo-Hookean residual and consistent tangent, one wavefront per element.
// Replaces autograd of NeoHookeanModel.energy (builtin.py:301-333,
// element_assembler.py:631-657): psi = mu/2 (I1 - d) - mu ln J + lam/2 (ln J)^2
// with F = I + grad u, grad u[a][d] = sum_b u[n_b][a] dphi_b/dx_d.
//   P   = mu F + (lam ln J - mu) F^-T
//   R_i[a]       = sum_q jxw_q sum_A P_q[a][A] g_q,i[A]
//   K_ij[a][b]   = sum_q jxw_q ( mu delta_ab g_i.g_j + lam h_i[a] h_j[b] - c2 h_i[b] h_j[a] )
// with h_i = F^-T g_i  (h_i[a] = sum_A g_i[A] Finv[A][a]) and c2 = lam ln J - mu:
// the 9x9 tangent dP/dF = mu I + lam F^-T (x) F^-T - c2 (F^-1 swapped) contracted
// with the shape gradients on both sides.  At u = 0 it equals the linear
// elasticity block.  Tangent layout as the block forms: ((e*B+i)*B+j)*D*D + a*D + b;
// residual slots [e][i][a] summed into [N][D] by the node incidence reduction.
typedef float tm_floatx16 __attribute__((ext_vector_type(16)));

// MQ: largest number of quadrature points the LDS arrays hold.  Slot row
// (e, i) (B blocks of D x D values) goes to e*B + i, or with inc_inv to its
// position inc_inv[e*B + i] in the node-sorted incidence list.  The VALU
// tangent leaves from registers: each lane stores its D x D block (72
// contiguous bytes at f64, lanes j, j+1 of a row side by side), which needs no
// LDS staging and measured faster than staged 16-byte row stores (82 instead
// of 130 VGPRs).  The MFMA tile spreads a row over lanes, so that variant is
// staged in LDS and written as whole slot rows with vector stores.
// upper (TM_FLAG_UPPER): only the pairs i <= j.
// EPW elements per wavefront, each on L = 64 / EPW lanes: the VALU path packs
// two, so that the per-quadrature-point stages (Q active lanes) serve two
// elements at once (tetra10 f64 1.15 -> 1.07 ms per Newton assembly).
// BIG (tetra20, hexahedron27: element tangents of 60 x 60 and 81 x 81): one
// wavefront per workgroup and the rule tables read through the caches instead
// of LDS, so that a wave needs LDS only for its gradients (2 Q B D doubles).
template <typename T, int D, int B, bool MFMA, int MQ, bool BIG, int EPW, bool TAN>
__global__ void __launch_bounds__(kWave * (BIG ? 1 : kWavesPerBlock))
    hyper_wave_kernel(const int32_t* __restrict__ conn, const T* __restrict__ pts, const T* __restrict__ disp,
                      const double* __restrict__ tab, int Q, double lam, double mu, int64_t n_elem, int upper,
                      T* __restrict__ tslot, T* __restrict__ rslot, const int32_t* __restrict__ inc_inv,
                      double* __restrict__ energy_e, const int32_t* __restrict__ slot_dest) {
  // TAN: the launch forms the tangent (ts == tslot); energy and residual
  // launches compile without the h arrays, the tangent stage and its
  // registers
  T* const ts = TAN ? tslot : nullptr;
  constexpr int kMaxQ = MQ;
  constexpr int kRowVals = B * D * D;  // values of one slot row
  constexpr int kWpb = BIG ? 1 : kWavesPerBlock;  // waves per workgroup
  static_assert(!(BIG && MFMA), "the MFMA tile holds at most 32 x 32");
  static_assert(EPW == 1 || !(MFMA || BIG), "the MFMA tile and BIG cells take a whole wavefront");
  constexpr int L = kWave / EPW;  // lanes per element
  constexpr int kEpb = kWpb * EPW;  // elements per workgroup
  constexpr bool kStage = MFMA;  // the MFMA tile spreads a row over lanes: staged row stores
  __shared__ double s_tab[BIG ? 1 : MQ * (1 + B + D * B)];
  __shared__ double s_x[kEpb][B * D];
  __shared__ double s_u[kEpb][B * D];
  __shared__ double s_inv[kEpb][kMaxQ][D * D];
  __shared__ double s_jxw[kEpb][kMaxQ];
  __shared__ double s_g[kEpb][kMaxQ * B * D];
  __shared__ double s_h[kEpb][TAN ? kMaxQ * B * D : 1];
  __shared__ double s_P[kEpb][kMaxQ][D * D];
  __shared__ double s_c2[kEpb][kMaxQ];
  __shared__ double s_psi[kEpb][kMaxQ];   // psi * jxw per point (energy launches)
  __shared__ double s_Fi[kEpb][kMaxQ][D * D];
  __shared__ int s_v[kEpb][B];
  __shared__ T s_out[kEpb][kStage ? B * kRowVals : 1];
  __shared__ uint16_t s_pair[B * B];

  static_assert(B <= L && MQ <= L, "an element's lanes index its nodes and its quadrature points");
  const int lane = threadIdx.x % L;   // lane within the element's group
  const int wv = threadIdx.x / L;     // element slot within the workgroup
  const int64_t e = (int64_t)blockIdx.x * kEpb + wv;
  const bool active = e < n_elem;
  const int tab_len = Q + Q * B + Q * D * B;
  if constexpr (!BIG)
    for (int t = threadIdx.x; t < tab_len; t += blockDim.x) s_tab[t] = tab[t];
  const int n_pairs = tmh::fill_pairs<B>(s_pair, upper != 0);
  const double* tb = BIG ? tab : s_tab;
  const double* w = tb;
  const double* dphi = tb + Q + Q * B;
  if (active && lane < B) s_v[wv][lane] = conn[e * B + lane];
  // the lane's tangent block destinations (contribution order), loaded now
  // so that their latency hides behind the per-point stages
  constexpr int kPairRounds = (B * B + L - 1) / L;
  int32_t dst[kPairRounds];
#pragma unroll
  for (int r = 0; r < kPairRounds; ++r) {
    const int p = lane + r * L;
    dst[r] = (!MFMA && slot_dest && ts && active && p < n_pairs) ? slot_dest[e * n_pairs + p] : 0;
  }
  __syncthreads();
  if (active) {
    for (int t = lane; t < B * D; t += L) {
      const int b = t / D, d = t % D;
      s_x[wv][t] = (double)pts[(int64_t)s_v[wv][b] * D + d];
      s_u[wv][t] = (double)disp[(int64_t)s_v[wv][b] * D + d];
    }
  }
  __syncthreads();
  // the Jacobians J[i][j] = sum_b dphi_b[i] x_b[j] (into s_inv, which
  // point q's lane then overwrites with the inverse)
  if (active)
    tmh::point_sums<D, B, L, MQ>(
        lane, Q, false, [&](int q, int i, int b) { return dphi[(q * D + i) * B + b]; },
        [&](int, int b, int j) { return s_x[wv][b * D + j]; },
        [&](int q, int i, int j) -> double& { return s_inv[wv][q][i * D + j]; });
  __syncthreads();
  if (active && lane < Q) {
    const int q = lane;
    double J[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) J[i][j] = s_inv[wv][q][i * D + j];
    double inv[D][D];
    double det;
    if constexpr (D == 2) {
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      inv[0][0] = J[1][1] / det;
      inv[0][1] = -J[0][1] / det;
      inv[1][0] = -J[1][0] / det;
      inv[1][1] = J[0][0] / det;
    } else {
      const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
      const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      const double r = 1.0 / det;
      inv[0][0] = c00 * r;
      inv[1][0] = c01 * r;
      inv[2][0] = c02 * r;
      inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r;
      inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r;
      inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r;
      inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
      inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r;
      inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
    }
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) s_inv[wv][q][i * D + j] = inv[i][j];
    s_jxw[wv][q] = w[q] * fabs(det);
  }
  __syncthreads();
  if (active) {
    for (int t = lane; t < Q * B * D; t += L) {
      const int q = t / (B * D), b = (t / D) % B, d = t % D;
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < D; ++i) s += dphi[(q * D + i) * B + b] * s_inv[wv][q][d * D + i];
      s_g[wv][t] = s;
    }
  }
  __syncthreads();
  // the deformation gradients F[a][d] = delta_ad + sum_b u_b[a] g_b[d] (into
  // s_Fi, which the per-point stage below overwrites with F^-1)
  if (active)
    tmh::point_sums<D, B, L, MQ>(
        lane, Q, true, [&](int, int a, int b) { return s_u[wv][b * D + a]; },
        [&](int q, int b, int d) { return s_g[wv][(q * B + b) * D + d]; },
        [&](int q, int a, int d) -> double& { return s_Fi[wv][q][a * D + d]; });
  __syncthreads();
  if (active && lane < Q) {
    const int q = lane;
    double F[D][D];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int d = 0; d < D; ++d) F[a][d] = s_Fi[wv][q][a * D + d];
    double Fi[D][D];
    double Jd;
    if constexpr (D == 2) {
      Jd = F[0][0] * F[1][1] - F[0][1] * F[1][0];
      Fi[0][0] = F[1][1] / Jd;
      Fi[0][1] = -F[0][1] / Jd;
      Fi[1][0] = -F[1][0] / Jd;
      Fi[1][1] = F[0][0] / Jd;
    } else {
      const double c00 = F[1][1] * F[2][2] - F[1][2] * F[2][1];
      const double c01 = F[1][2] * F[2][0] - F[1][0] * F[2][2];
      const double c02 = F[1][0] * F[2][1] - F[1][1] * F[2][0];
      Jd = F[0][0] * c00 + F[0][1] * c01 + F[0][2] * c02;
      const double r = 1.0 / Jd;
      Fi[0][0] = c00 * r;
      Fi[1][0] = c01 * r;
      Fi[2][0] = c02 * r;
      Fi[0][1] = (F[0][2] * F[2][1] - F[0][1] * F[2][2]) * r;
      Fi[1][1] = (F[0][0] * F[2][2] - F[0][2] * F[2][0]) * r;
      Fi[2][1] = (F[0][1] * F[2][0] - F[0][0] * F[2][1]) * r;
      Fi[0][2] = (F[0][1] * F[1][2] - F[0][2] * F[1][1]) * r;
      Fi[1][2] = (F[0][2] * F[1][0] - F[0][0] * F[1][2]) * r;
      Fi[2][2] = (F[0][0] * F[1][1] - F[0][1] * F[1][0]) * r;
    }
    const double lj = log(Jd);
    const double c2 = lam * lj - mu;
    s_c2[wv][q] = c2;
    if (energy_e) {
      // psi = mu/2 (I1 - d) - mu ln J + lam/2 (ln J)^2, I1 = sum F_aA^2 (builtin.py:301-314)
      double I1 = 0.0;
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int A = 0; A < D; ++A) I1 += F[a][A] * F[a][A];
      s_psi[wv][q] = (0.5 * mu * (I1 - (double)D) - mu * lj + 0.5 * lam * lj * lj) * s_jxw[wv][q];
    }
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int A = 0; A < D; ++A) {
        s_P[wv][q][a * D + A] = mu * F[a][A] + c2 * Fi[A][a];
        s_Fi[wv][q][a * D + A] = Fi[a][A];
      }
  }
  __syncthreads();
  if (energy_e && active && lane == 0) {
    // the element's energy, its points summed in order
    double s = 0.0;
    for (int q = 0; q < Q; ++q) s += s_psi[wv][q];
    energy_e[e] = s;
  }
  if (active && ts) {
    // h = F^-T g, read by the tangent only
    for (int t = lane; t < Q * B * D; t += L) {
      const int q = t / (B * D), b = (t / D) % B, a = t % D;
      const double* g = &s_g[wv][(q * B + b) * D];
      double s = 0.0;
#pragma unroll
      for (int A = 0; A < D; ++A) s += g[A] * s_Fi[wv][q][A * D + a];
      s_h[wv][t] = s;
    }
  }
  __syncthreads();
  // (no early return for inactive waves: the block meets once more below)
  if (rslot && active) {
    for (int t = lane; t < B * D; t += L) {
      const int i = t / D, a = t % D;
      double acc = 0.0;
      for (int q = 0; q < Q; ++q) {
        const double* g = &s_g[wv][(q * B + i) * D];
        double s = 0.0;
#pragma unroll
        for (int A = 0; A < D; ++A) s += s_P[wv][q][a * D + A] * g[A];
        acc += s * s_jxw[wv][q];
      }
      // element order: row (e, i) at e*B + i; incidence order: at inc_inv[e*B + i]
      const int64_t row = inc_inv ? (int64_t)inc_inv[e * B + i] : e * B + i;
      rslot[row * D + a] = (T)acc;
    }
  }
  if constexpr (MFMA) {
    // K_e = sum_q L_q^T R_q as one 32 x 32 tile (rows (i, a), columns (j, b),
    // B*D <= 32) with inner index (q, c, A), two per v_mfma_f32_32x32x2f32:
    //   L^T[(i,a)][(q,c,A)] = delta_ac g_q,i[A]
    //   R[(q,c,A)][(j,b)]   = jxw_q (mu delta_cb g_q,j[A] + lam Finv[A][c] h_j[b] - c2 Finv[A][b] h_j[c])
    // Lane l supplies A[m = l%32][k = l/32] and B[k = l/32][n = l%32]; the
    // accumulator register r of lane l holds row (r/4)*8 + (l/32)*4 + r%4,
    // column l%32.
    if (ts) {
      constexpr int DD = D * D;
      constexpr int M = B * D;
      static_assert(M <= 32, "tile holds at most 32 rows");
      const int m = lane % 32, kh = lane / 32;
      const int i = m / D, a = m % D;
      const int j = m / D, b = m % D;  // n == m for the B operand column
      tm_floatx16 acc;
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
      const int steps = (Q * DD) / 2;
      for (int st = 0; st < steps; ++st) {
        const int kk = 2 * st + kh;
        const int q = kk / DD, c = (kk % DD) / D, A = kk % D;
        float av = 0.0f, bv = 0.0f;
        if (m < M) {
          const double* g = &s_g[wv][q * B * D];
          const double* h = &s_h[wv][q * B * D];
          const double* Fi = s_Fi[wv][q];
          if (a == c) av = (float)g[i * D + A];
          double v = lam * Fi[A * D + c] * h[j * D + b] - s_c2[wv][q] * Fi[A * D + b] * h[j * D + c];
          if (c == b) v += mu * g[j * D + A];
          bv = (float)(v * s_jxw[wv][q]);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
      }
      const int n = m;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = (r / 4) * 8 + kh * 4 + (r % 4);
        if (row < M && n < M && active) {
          const int ii = row / D, aa = row % D, jj = n / D, bb = n % D;
          s_out[wv][(ii * B + jj) * DD + aa * D + bb] = (T)acc[r];
        }
      }
    }
  } else if (ts && active) {
    // one lane per node pair (i, j): the whole D x D block from one read of
    // g_i, g_j, h_i, h_j per quadrature point
    constexpr int DD = D * D;
#pragma unroll
    for (int r = 0; r < kPairRounds; ++r) {
      const int p = lane + r * L;
      if (p >= n_pairs) break;
      int i = s_pair[p] >> 8, j = s_pair[p] & 255;
      T* out;
      if constexpr (!kStage) {
        // element order: block p of the element's pairs (packed when upper);
        // incidence order (full blocks only): slot row inc_inv[e*B + i];
        // contribution order (upper): slot slot_dest[e*P + p], transposed
        // when its id is negative -- formed as block (j, i), which is the
        // exact transpose of (i, j) (the same products, commuted)
        if (slot_dest) {
          const int32_t d = dst[r];
          out = ts + (int64_t)(d & INT32_MAX) * DD;
          if (d < 0) {
            const int t = i;
            i = j;
            j = t;
          }
        } else {
          out = inc_inv ? ts + (int64_t)inc_inv[e * B + i] * kRowVals + j * DD
                        : ts + (e * n_pairs + p) * DD;
        }
      } else {
        out = &s_out[wv][(i * B + j) * DD];
      }
      double acc[DD];
#pragma unroll
      for (int k = 0; k < DD; ++k) acc[k] = 0.0;
      for (int q = 0; q < Q; ++q) {
        const double* gi = &s_g[wv][(q * B + i) * D];
        const double* gj = &s_g[wv][(q * B + j) * D];
        const double* hi = &s_h[wv][(q * B + i) * D];
        const double* hj = &s_h[wv][(q * B + j) * D];
        double vgi[D], vgj[D], vhi[D], vhj[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
          vgi[d] = gi[d];
          vgj[d] = gj[d];
          vhi[d] = hi[d];
          vhj[d] = hj[d];
        }
        const double w = s_jxw[wv][q];
        const double lw = lam * w, cw = s_c2[wv][q] * w;
        double dot = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) dot += vgi[d] * vgj[d];
        const double mw = mu * dot * w;
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
          for (int b = 0; b < D; ++b) {
            double t = lw * (vhi[a] * vhj[b]) - cw * (vhi[b] * vhj[a]);
            if (a == b) t += mw;
            acc[a * D + b] += t;
          }
      }
#pragma unroll
      for (int k = 0; k < DD; ++k) out[k] = (T)acc[k];
    }
  }
  if constexpr (kStage) {
    __syncthreads();
    if (ts && active) {
      // whole slot rows with vector stores: 16 bytes when a row's size allows,
      // else 8 (slot buffers are 256-byte aligned)
      constexpr int kRowBytes = kRowVals * (int)sizeof(T);
      constexpr int kVec = kRowBytes % 16 == 0 ? 16 : 8;
      static_assert(kRowBytes % kVec == 0, "slot rows are whole vectors");
      constexpr int kPerRow = kRowBytes / kVec;
      typedef int32_t vec_t __attribute__((ext_vector_type(kVec / 4)));
      for (int v = lane; v < B * kPerRow; v += L) {
        const int i = v / kPerRow, o = v % kPerRow;
        const int64_t srow = inc_inv ? (int64_t)inc_inv[e * B + i] : e * B + i;
        const vec_t x = reinterpret_cast<const vec_t*>(&s_out[wv][i * kRowVals])[o];
        reinterpret_cast<vec_t*>(ts + srow * kRowVals)[o] = x;
      }
    }
  }
}

template <typename T, int D, int B>
int launch_hyper(const tm_mesh_view* m, const tm_form_desc* f, const void* disp, void* tslot, void* rslot,
                 const int32_t* inc_inv, double* energy_e, const int32_t* slot_dest, hipStream_t stream) {
  // more than 32 element dofs (tetra20: 60, hexahedron27: 81), or a rule of
  // more than 16 points (the 24-point tetrahedron rule): one element per
  // workgroup, the BIG variant, with rules of up to 32 points
  constexpr bool kBigCell = B * D > 32;
  const int Q = f->n_quad;
  const bool big = kBigCell || Q > kMaxQ;
  const int wpb = big ? 1 : kWavesPerBlock;
  if (Q <= 0 || Q > kMaxQWave) return tmh::fail(TM_ERR_UNSUPPORTED, "quadrature rule has too many points");
  if (!f->tab_dev) return tmh::fail(TM_ERR_ARG, "device form tables missing");
  const double mu = f->young / (2.0 * (1.0 + f->poisson));
  const double lam = f->young * f->poisson / ((1.0 + f->poisson) * (1.0 - 2.0 * f->poisson));
  const bool mfma = (f->flags & TM_FLAG_MFMA) != 0;
  if (mfma && !(std::is_same<T, float>::value && B * D <= 32))
    return tmh::fail(TM_ERR_UNSUPPORTED, "the MFMA tangent path is float32 only");
  if (mfma && Q > kMaxQ) return tmh::fail(TM_ERR_UNSUPPORTED, "the MFMA tangent path takes rules of <= 16 points");
  if (mfma && (Q * D * D) % 2 != 0) return tmh::fail(TM_ERR_UNSUPPORTED, "MFMA path needs an even Q*D*D");
  // two elements per wavefront (32 lanes each) on the VALU path by default:
  // the per-quadrature-point stages, which use Q lanes, then serve two
  // elements; four (16 lanes each) without a tangent, whose pair loop wants
  // no lanes (tetra10 energy + residual 185 -> 146 us per call, f64); form->epw
  // asks for 1, 2 or 4 (4: 16 lanes, rules of <= 4 points)
  int epw = (mfma || big) ? 1 : (!tslot && Q <= 4 && B <= 16) ? 4 : 2;
  if (f->epw != 0) {
    const bool ok = f->epw == 1 || (!mfma && !big && (f->epw == 2 || (f->epw == 4 && Q <= 4 && B <= 16)));
    if (!ok) return tmh::fail(TM_ERR_ARG, "elements per wavefront not admitted for this cell / rule / tangent");
    epw = f->epw;
  }
  const int64_t per_block = (int64_t)wpb * epw;
  const int64_t grid = (m->n_elem + per_block - 1) / per_block;
  if (grid == 0) return TM_OK;
  const dim3 g((unsigned)grid), blk(kWave * wpb);
  const int upper = (f->flags & TM_FLAG_UPPER) != 0;
  if (upper && (mfma || inc_inv))
    return tmh::fail(TM_ERR_ARG, "TM_FLAG_UPPER needs the element-ordered VALU tangent (no MFMA, no block rows)");
#define TM_HYPER_T(MF, MQ, BG, EPW, TAN)                                                            \
  hyper_wave_kernel<T, D, B, MF, MQ, BG, EPW, TAN><<<g, blk, 0, stream>>>(                          \
      m->conn, (const T*)m->points, (const T*)disp, f->tab_dev, Q, lam, mu, m->n_elem, upper,     \
      (T*)tslot, (T*)rslot, inc_inv, energy_e, upper ? slot_dest : nullptr)
#define TM_HYPER_V(MF, MQ, BG, EPW) TM_HYPER_T(MF, MQ, BG, EPW, true)
  // the VALU instances come without the tangent machinery for energy and
  // residual launches (fewer registers and LDS: more waves per SIMD)
#define TM_HYPER(MF, MQ, EPW)                                \
  do {                                                       \
    if constexpr (MF) TM_HYPER_T(MF, MQ, false, EPW, true);  \
    else if (tslot) TM_HYPER_T(MF, MQ, false, EPW, true);    \
    else TM_HYPER_T(MF, MQ, false, EPW, false);              \
  } while (0)
  if constexpr (kBigCell) {
    if (Q <= 16) TM_HYPER_V(false, 16, true, 1); else TM_HYPER_V(false, 32, true, 1);
  } else if (big) {
    TM_HYPER_V(false, 32, true, 1);
  } else if (mfma) {
    if constexpr (std::is_same<T, float>::value) {  // float64 was rejected above
      if (Q <= 4) TM_HYPER(true, 4, 1); else TM_HYPER(true, 16, 1);
    }
  } else if (Q <= 4) {
    if (epw == 1) TM_HYPER(false, 4, 1);
    else if (epw == 2) TM_HYPER(false, 4, 2);
    else if constexpr (B <= 16) TM_HYPER(false, 4, 4);
  } else {
    if (epw == 1) TM_HYPER(false, 16, 1); else TM_HYPER(false, 16, 2);
  }
#undef TM_HYPER
#undef TM_HYPER_V
#undef TM_HYPER_T
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("neo-hookean kernel launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
}

template <typename T>
int dispatch_hyper(const tm_mesh_view* m, const tm_form_desc* f, const void* disp, void* tslot, void* rslot,
                   const int32_t* inc_inv, double* energy_e, hipStream_t s, const int32_t* slot_dest = nullptr) {
  switch (m->cell) {
#define TM_CASE(C, D, B) \
  case C: return launch_hyper<T, D, B>(m, f, disp, tslot, rslot, inc_inv, energy_e, slot_dest, s);
    TM_ALL_CELLS(TM_CASE)
#undef TM_CASE
    default: return tmh::fail(TM_ERR_ARG, "unknown cell");
  }
}

template <typename T, int D, int B, int MQ>
int launch_wave(const tm_mesh_view* m, const tm_form_desc* f, void* slot_ws, hipStream_t stream, int* hh_out,
                const int32_t* slot_dest) {
  const int Q = f->n_quad;
  const T* pts = static_cast<const T*>(m->points);
  const T* coeff = static_cast<const T*>(f->coeff);
  T* slot = static_cast<T*>(slot_ws);
  // two elements per wavefront (32 lanes each) where their LDS arrays fit a
  // 40 KB workgroup: the per-quadrature-point stage (Q active lanes) then
  // serves two elements
  constexpr int kElemBytes = 8 * (B * D + MQ * (D * D + 2 + B * D)) + 4 * B;
  constexpr int kEpw = (2 * kWavesPerBlock * kElemBytes <= 40960 && MQ <= kWave / 2 && B <= kWave / 2) ? 2 : 1;
  // form->epw: 1, or 2 where two elements' arrays fit (0: the largest that fits)
  const int epw = f->epw == 0 ? kEpw : f->epw;
  if (epw != 1 && epw != kEpw) return tmh::fail(TM_ERR_ARG, "elements per wavefront not admitted for this cell / rule");
  const int64_t per_block = (int64_t)kWavesPerBlock * epw;
  const int64_t grid = (m->n_elem + per_block - 1) / per_block;
  if (grid == 0) return TM_OK;
  const dim3 block(kWave * kWavesPerBlock);
  double lam = 0.0, mu = 0.0;
  *hh_out = 1;
  const int upper = (f->flags & TM_FLAG_UPPER) != 0;
#define TM_WAVE(H, FORM)                                                                                          \
  do {                                                                                                            \
    if (epw == kEpw)                                                                                              \
      wave_element_kernel<T, D, B, H, FORM, MQ, kEpw><<<grid, block, 0, stream>>>(                               \
          m->conn, pts, coeff, f->tab_dev, Q, lam, mu, m->n_elem, upper, slot, upper ? slot_dest : nullptr);   \
    else                                                                                                          \
      wave_element_kernel<T, D, B, H, FORM, MQ, 1><<<grid, block, 0, stream>>>(                                  \
          m->conn, pts, coeff, f->tab_dev, Q, lam, mu, m->n_elem, upper, slot, upper ? slot_dest : nullptr);   \
  } while (0)
  switch (f->form) {
    case TM_FORM_LAPLACE:
      TM_WAVE(1, TM_FORM_LAPLACE);
      break;
    case TM_FORM_DIFFUSION_ELEM:
      if (!coeff) return tmh::fail(TM_ERR_ARG, "diffusion form needs a coefficient");
      TM_WAVE(1, TM_FORM_DIFFUSION_ELEM);
      break;
    case TM_FORM_DIFFUSION_NODE:
      if (!coeff) return tmh::fail(TM_ERR_ARG, "diffusion form needs a coefficient");
      TM_WAVE(1, TM_FORM_DIFFUSION_NODE);
      break;
    case TM_FORM_MASS:
      TM_WAVE(1, TM_FORM_MASS);
      break;
    case TM_FORM_LINEAR_ELASTICITY:
      mu = f->young / (2.0 * (1.0 + f->poisson));
      lam = f->young * f->poisson / ((1.0 + f->poisson) * (1.0 - 2.0 * f->poisson));
      *hh_out = D * D;
      TM_WAVE(D, TM_FORM_LINEAR_ELASTICITY);
      break;
    default:
      return tmh::fail(TM_ERR_UNSUPPORTED, "unknown form");
  }
#undef TM_WAVE
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("element kernel launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
}

template <typename T, int D, int B>
int launch_scalar_or_block(const tm_mesh_view* m, const tm_form_desc* f, void* slot_ws, hipStream_t stream,
                           int* hh_out, const int32_t* sd) {
  const int Q = f->n_quad;
  if (Q <= 0 || Q > kMaxQWave) return tmh::fail(TM_ERR_UNSUPPORTED, "quadrature rule has too many points");
  if (!f->tab_dev) return tmh::fail(TM_ERR_ARG, "device form tables missing");
  if (Q <= 4) return launch_wave<T, D, B, 4>(m, f, slot_ws, stream, hh_out, sd);
  if (Q <= 8) return launch_wave<T, D, B, 8>(m, f, slot_ws, stream, hh_out, sd);
  if (Q <= 16) return launch_wave<T, D, B, 16>(m, f, slot_ws, stream, hh_out, sd);
  return launch_wave<T, D, B, kMaxQWave>(m, f, slot_ws, stream, hh_out, sd);
}

template <typename T>
int dispatch_cell(const tm_mesh_view* m, const tm_form_desc* f, void* slot_ws, hipStream_t stream, int* hh,
                  const int32_t* sd) {
  switch (m->cell) {
#define TM_CASE(C, D, B) \
  case C: return launch_scalar_or_block<T, D, B>(m, f, slot_ws, stream, hh, sd);
    TM_ALL_CELLS(TM_CASE)
#undef TM_CASE
    default: return tmh::fail(TM_ERR_ARG, "unknown cell");
  }
}

}  // namespace

int tm_hyper_element_rows(const tm_mesh_view* mesh, const tm_form_desc* form, const void* disp, void* tslot,
                          void* rslot, const int32_t* inc_inv, hipStream_t stream) {
  return mesh->dtype == TM_F64 ? dispatch_hyper<double>(mesh, form, disp, tslot, rslot, inc_inv, nullptr, stream)
                               : dispatch_hyper<float>(mesh, form, disp, tslot, rslot, inc_inv, nullptr, stream);
}

int tm_p2_element_matrices(const tm_mesh_view* mesh, const tm_form_desc* form, void* slot_ws,
                           hipStream_t stream, int* hh_out, const int32_t* slot_dest) {
  if (mesh->dtype == TM_F64) return dispatch_cell<double>(mesh, form, slot_ws, stream, hh_out, slot_dest);
  return dispatch_cell<float>(mesh, form, slot_ws, stream, hh_out, slot_dest);
}

extern "C" int tm_assemble_neohookean(const tm_mesh_view* mesh, const tm_form_desc* form, const void* disp,
                                      const tm_pattern_view* pat, const tm_incidence_view* inc, void* tslot_ws,
                                      void* tangent, void* rslot_ws, void* residual, tm_stream_t stream_) {
  TM_TRY_BEGIN
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  if (!mesh || !form || !disp) return tmh::fail(TM_ERR_ARG, "NULL argument");
  if (tmh::cell_basis(mesh->cell) < 0) return tmh::fail(TM_ERR_ARG, "unknown cell");
  const int64_t B = tmh::cell_basis(mesh->cell), D = tmh::cell_dim(mesh->cell);
  if (mesh->n_elem * B * B >= tmh::kMaxIndex) return tmh::fail(TM_ERR_SIZE, "mesh too large");
  if (tangent && (!pat || !tslot_ws)) return tmh::fail(TM_ERR_ARG, "tangent needs a pattern and a slot workspace");
  if (residual && (!inc || !rslot_ws)) return tmh::fail(TM_ERR_ARG, "residual needs incidence and a workspace");
  void* ts = tangent ? tslot_ws : nullptr;
  void* rs = residual ? rslot_ws : nullptr;
  if (tangent && pat->slot_dest && !(form->flags & TM_FLAG_UPPER))
    return tmh::fail(TM_ERR_ARG, "slot_dest needs TM_FLAG_UPPER (symmetric element blocks)");
  const int32_t* sd = tangent ? pat->slot_dest : nullptr;
  int st = mesh->dtype == TM_F64 ? dispatch_hyper<double>(mesh, form, disp, ts, rs, nullptr, nullptr, stream, sd)
                                 : dispatch_hyper<float>(mesh, form, disp, ts, rs, nullptr, nullptr, stream, sd);
  if (st) return st;
  // the residual reduction first: part of its slots (tens of MB) still sits in
  // the caches after the element kernel, and would not after the tangent
  // reduction has streamed the much larger tangent slots (measured faster)
  if (residual) {
    st = tm_segmented_reduce(mesh->dtype, mesh->n_points, (int32_t)D, inc->inc_ptr, inc->inc, inc->n_blocks,
                             inc->blk_ptr, rslot_ws, residual, stream_);
    if (st) return st;
  }
  if (tangent) {
    st = tmh::reduce_view(mesh->dtype, pat, (int32_t)(D * D), tslot_ws, tangent, stream_);
    if (st) return st;
  }
  return TM_OK;
  TM_TRY_END
}

extern "C" int tm_neohookean_energy(const tm_mesh_view* mesh, const tm_form_desc* form, const void* disp,
                                    void* energy_ws, void* energy, const tm_incidence_view* inc, void* rslot_ws,
                                    void* residual, tm_stream_t stream_) {
  TM_TRY_BEGIN
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  if (!mesh || !form || !disp || !energy || !energy_ws) return tmh::fail(TM_ERR_ARG, "NULL argument");
  if (mesh->dtype != TM_F64 && mesh->dtype != TM_F32) return tmh::fail(TM_ERR_ARG, "unknown dtype");
  const int64_t B = tmh::cell_basis(mesh->cell), D = tmh::cell_dim(mesh->cell);
  if (B < 0) return tmh::fail(TM_ERR_ARG, "unknown cell");
  if (mesh->n_elem < 0 || mesh->n_elem * B * D >= tmh::kMaxIndex) return tmh::fail(TM_ERR_SIZE, "mesh too large");
  if (mesh->n_elem > 0 && (!mesh->conn || !mesh->points)) return tmh::fail(TM_ERR_ARG, "NULL buffer");
  if (residual && (!inc || !rslot_ws)) return tmh::fail(TM_ERR_ARG, "residual needs incidence and a workspace");
  // energy launches form no tangent: the tangent flags do not apply
  tm_form_desc f = *form;
  f.flags = 0;
  double* e_out = static_cast<double*>(energy_ws);
  void* rs = residual ? rslot_ws : nullptr;
  int st = mesh->dtype == TM_F64 ? dispatch_hyper<double>(mesh, &f, disp, nullptr, rs, nullptr, e_out, stream)
                                 : dispatch_hyper<float>(mesh, &f, disp, nullptr, rs, nullptr, e_out, stream);
  if (st) return st;
  st = tmh::sum_fixed(e_out, mesh->n_elem, e_out + mesh->n_elem, mesh->dtype, energy, stream_);
  if (st) return st;
  if (residual) {
    st = tm_segmented_reduce(mesh->dtype, mesh->n_points, (int32_t)D, inc->inc_ptr, inc->inc, inc->n_blocks,
                             inc->blk_ptr, rslot_ws, residual, stream_);
    if (st) return st;
  }
  return TM_OK;
  TM_TRY_END
}
