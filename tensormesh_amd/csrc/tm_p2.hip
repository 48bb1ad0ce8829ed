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

#include "tm_internal.h"

namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kMaxQ = 16;
constexpr int kMaxTab = kMaxQ + kMaxQ * 10 + kMaxQ * 3 * 10;

template <typename T, int D, int B, int H, int FORM>
__global__ void __launch_bounds__(kWave * kWavesPerBlock)
    wave_element_kernel(const int32_t* __restrict__ conn, const T* __restrict__ pts,
                        const T* __restrict__ coeff, const double* __restrict__ tab, int Q,
                        double lam, double mu, int64_t n_elem, T* __restrict__ slot) {
  __shared__ double s_tab[kMaxTab];
  __shared__ double s_x[kWavesPerBlock][B * D];
  __shared__ double s_inv[kWavesPerBlock][kMaxQ][D * D];
  __shared__ double s_jxw[kWavesPerBlock][kMaxQ];
  __shared__ double s_kq[kWavesPerBlock][kMaxQ];
  __shared__ double s_g[kWavesPerBlock][kMaxQ * B * D];
  __shared__ int s_v[kWavesPerBlock][B];

  const int lane = threadIdx.x % kWave;
  const int wv = threadIdx.x / kWave;
  const int64_t e = (int64_t)blockIdx.x * kWavesPerBlock + wv;
  const bool active = e < n_elem;

  const int tab_len = Q + Q * B + Q * D * B;
  for (int t = threadIdx.x; t < tab_len; t += blockDim.x) s_tab[t] = tab[t];
  const double* w = s_tab;
  const double* phi = s_tab + Q;
  const double* dphi = s_tab + Q + Q * B;

  if (active) {
    if (lane < B) s_v[wv][lane] = conn[e * B + lane];
  }
  __syncthreads();
  if (active && lane < B * D) {
    const int b = lane / D, d = lane % D;
    s_x[wv][lane] = (double)pts[(int64_t)s_v[wv][b] * D + d];
  }
  __syncthreads();
  if (active && lane < Q) {
    const int q = lane;
    double J[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        double s = 0.0;
        for (int b = 0; b < B; ++b) s += s_x[wv][b * D + j] * dphi[(q * D + i) * B + b];
        J[i][j] = s;
      }
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
    for (int t = lane; t < Q * B * D; t += kWave) {
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
  T* out = slot + e * (int64_t)(B * B * HH);
  for (int t = lane; t < B * B * HH; t += kWave) {
    const int ij = t / HH, ac = t % HH;
    const int i = ij / B, j = ij % B, a = ac / H, c = ac % H;
    double acc = 0.0;
    for (int q = 0; q < Q; ++q) {
      const double* g = &s_g[wv][q * B * D];
      double val;
      if constexpr (FORM == TM_FORM_MASS) {
        val = phi[q * B + i] * phi[q * B + j];
      } else if constexpr (FORM == TM_FORM_LINEAR_ELASTICITY) {
        double dot = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) dot += g[i * D + d] * g[j * D + d];
        val = lam * g[i * D + a] * g[j * D + c] + mu * g[i * D + c] * g[j * D + a] + (a == c ? mu * dot : 0.0);
      } else {
        double dot = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) dot += g[i * D + d] * g[j * D + d];
        val = s_kq[wv][q] * dot;
      }
      acc += val * s_jxw[wv][q];
    }
    out[t] = (T)acc;
  }
}

template <typename T, int D, int B>
int launch_scalar_or_block(const tm_mesh_view* m, const tm_form_desc* f, void* slot_ws, hipStream_t stream,
                           int* hh_out) {
  const int Q = f->n_quad;
  if (Q <= 0 || Q > kMaxQ) return tmh::fail(TM_ERR_UNSUPPORTED, "quadrature rule has too many points");
  if (!f->tab_dev) return tmh::fail(TM_ERR_ARG, "device form tables missing");
  const T* pts = static_cast<const T*>(m->points);
  const T* coeff = static_cast<const T*>(f->coeff);
  T* slot = static_cast<T*>(slot_ws);
  const int64_t grid = (m->n_elem + kWavesPerBlock - 1) / kWavesPerBlock;
  if (grid == 0) return TM_OK;
  const dim3 block(kWave * kWavesPerBlock);
  double lam = 0.0, mu = 0.0;
  *hh_out = 1;
  switch (f->form) {
    case TM_FORM_LAPLACE:
      wave_element_kernel<T, D, B, 1, TM_FORM_LAPLACE><<<grid, block, 0, stream>>>(m->conn, pts, coeff, f->tab_dev, Q, lam, mu, m->n_elem, slot);
      break;
    case TM_FORM_DIFFUSION_ELEM:
      if (!coeff) return tmh::fail(TM_ERR_ARG, "diffusion form needs a coefficient");
      wave_element_kernel<T, D, B, 1, TM_FORM_DIFFUSION_ELEM><<<grid, block, 0, stream>>>(m->conn, pts, coeff, f->tab_dev, Q, lam, mu, m->n_elem, slot);
      break;
    case TM_FORM_DIFFUSION_NODE:
      if (!coeff) return tmh::fail(TM_ERR_ARG, "diffusion form needs a coefficient");
      wave_element_kernel<T, D, B, 1, TM_FORM_DIFFUSION_NODE><<<grid, block, 0, stream>>>(m->conn, pts, coeff, f->tab_dev, Q, lam, mu, m->n_elem, slot);
      break;
    case TM_FORM_MASS:
      wave_element_kernel<T, D, B, 1, TM_FORM_MASS><<<grid, block, 0, stream>>>(m->conn, pts, coeff, f->tab_dev, Q, lam, mu, m->n_elem, slot);
      break;
    case TM_FORM_LINEAR_ELASTICITY:
      mu = f->young / (2.0 * (1.0 + f->poisson));
      lam = f->young * f->poisson / ((1.0 + f->poisson) * (1.0 - 2.0 * f->poisson));
      *hh_out = D * D;
      wave_element_kernel<T, D, B, D, TM_FORM_LINEAR_ELASTICITY><<<grid, block, 0, stream>>>(m->conn, pts, coeff, f->tab_dev, Q, lam, mu, m->n_elem, slot);
      break;
    default:
      return tmh::fail(TM_ERR_UNSUPPORTED, "unknown form");
  }
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("element kernel launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
}

template <typename T>
int dispatch_cell(const tm_mesh_view* m, const tm_form_desc* f, void* slot_ws, hipStream_t stream, int* hh) {
  switch (m->cell) {
    case TM_TRI3: return launch_scalar_or_block<T, 2, 3>(m, f, slot_ws, stream, hh);
    case TM_TET4: return launch_scalar_or_block<T, 3, 4>(m, f, slot_ws, stream, hh);
    case TM_TRI6: return launch_scalar_or_block<T, 2, 6>(m, f, slot_ws, stream, hh);
    case TM_TET10: return launch_scalar_or_block<T, 3, 10>(m, f, slot_ws, stream, hh);
    default: return tmh::fail(TM_ERR_ARG, "unknown cell");
  }
}

}  // namespace

int tm_p2_element_matrices(const tm_mesh_view* mesh, const tm_form_desc* form, void* slot_ws,
                           hipStream_t stream, int* hh_out) {
  if (mesh->dtype == TM_F64) return dispatch_cell<double>(mesh, form, slot_ws, stream, hh_out);
  return dispatch_cell<float>(mesh, form, slot_ws, stream, hh_out);
}
