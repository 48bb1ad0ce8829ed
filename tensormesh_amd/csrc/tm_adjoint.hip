// Hand-written adjoint of the P1 scalar assembly.
//
// The reference differentiates its assembly with autograd through
// index_add_, einsum, torch.inverse and torch.linalg.det
// (tensormesh/assemble/element_assembler.py:416-441, element.py:956-968,
// transformation.py:1265-1272); the cotangent on the matrix entries usually
// comes from the sparse solve's backward, dL/dA_ij = -lambda_i u_j
// (tensormesh/sparse/solve/torch_solve.py:262-272).  Here the same
// derivatives are computed in closed form, with the element geometry
// recomputed instead of stored.
//
// For a P1 simplex with edge matrix X = [x_1 - x_0, ..., x_D - x_0] (columns),
// G = X^-T (columns = grad(phi_1..D)), P = [-1 | I] and V = W |det X|
// (W = sum_q w_q), the local matrix is K_e = c_e V P^T G^T G P.  With the
// cotangent g_ij = dL/dK_e[i][j] (gathered through slot2nnz), S = sym(g) and
// Sh = P S P^T:
//   l_e        = c_e V tr(Sh G^T G)
//   dl/dc_e    = V tr(Sh G^T G)                      (t below)
//   dl/dX      = c_e V (t G - 2 G Sh G^T G)
//   dl/dx_k    = column k-1 of dl/dX,  dl/dx_0 = -sum_k dl/dx_k
// (d|det X| = |det X| tr(X^-1 dX), dG = -G dX^T G).  For the mass form
// K_e = |det X| Mref:  dl/dX = |det X| <g, Mref> G.
// Per-node sums run over the node incidence lists in ascending (e, i) order
// (the segmented reduction of tm_reduce.hip), so gradients are reproducible.
#include <hip/hip_runtime.h>

#include "tm_internal.h"

namespace {

constexpr int kBlock = 256;

struct AdjConsts {
  double W;
  double wphi[4];
  double mref[16];
};

template <typename T, int D, int FORM>
__global__ void __launch_bounds__(kBlock)
    p1_adjoint_kernel(const int32_t* __restrict__ conn, const T* __restrict__ pts, const T* __restrict__ coeff,
                      AdjConsts c, const int32_t* __restrict__ slot2nnz, const T* __restrict__ gvals,
                      int64_t n_elem, T* __restrict__ dcoef, T* __restrict__ gx, T* __restrict__ gcoef_slot) {
  constexpr int B = D + 1;
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= n_elem) return;
  int v[B];
#pragma unroll
  for (int b = 0; b < B; ++b) v[b] = conn[e * B + b];
  double x[B][D];
#pragma unroll
  for (int b = 0; b < B; ++b)
#pragma unroll
    for (int d = 0; d < D; ++d) x[b][d] = (double)pts[(int64_t)v[b] * D + d];
  // X columns = edges; G = X^-T: column k of G = cofactor vector n_k / det
  double Xc[D][D];  // Xc[k][d] = (x_{k+1} - x_0)[d]
#pragma unroll
  for (int k = 0; k < D; ++k)
#pragma unroll
    for (int d = 0; d < D; ++d) Xc[k][d] = x[k + 1][d] - x[0][d];
  double n[D][D];  // n[k][d]
  double det;
  if constexpr (D == 3) {
    const double* a = Xc[0];
    const double* bb = Xc[1];
    const double* cc = Xc[2];
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
    n[0][0] = Xc[1][1];  n[0][1] = -Xc[1][0];
    n[1][0] = -Xc[0][1]; n[1][1] = Xc[0][0];
    det = Xc[0][0] * Xc[1][1] - Xc[0][1] * Xc[1][0];
  }
  const double adet = fabs(det);
  const double rdet = 1.0 / det;
  double G[D][D];  // G[d][k] = n[k][d] / det   (row d = physical dim, column k)
#pragma unroll
  for (int d = 0; d < D; ++d)
#pragma unroll
    for (int k = 0; k < D; ++k) G[d][k] = n[k][d] * rdet;
  // cotangent of the local matrix
  double g[B][B];
  const int32_t* s2n = slot2nnz + e * (B * B);
#pragma unroll
  for (int i = 0; i < B; ++i)
#pragma unroll
    for (int j = 0; j < B; ++j) g[i][j] = (double)gvals[s2n[i * B + j]];

  double dX[D][D];  // dX[d][k] = dl / dXc[k][d]
  double dc = 0.0;
  if constexpr (FORM == TM_FORM_MASS) {
    double gm = 0.0;
#pragma unroll
    for (int i = 0; i < B; ++i)
#pragma unroll
      for (int j = 0; j < B; ++j) gm += g[i][j] * c.mref[i * B + j];
    const double f = adet * gm;
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
      for (int k = 0; k < D; ++k) dX[d][k] = f * G[d][k];
  } else {
    // Sh = P sym(g) P^T
    double Sh[D][D];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = 0; b < D; ++b) {
        const double s_ab = 0.5 * (g[a + 1][b + 1] + g[b + 1][a + 1]);
        const double s_a0 = 0.5 * (g[a + 1][0] + g[0][a + 1]);
        const double s_0b = 0.5 * (g[0][b + 1] + g[b + 1][0]);
        Sh[a][b] = g[0][0] - s_a0 - s_0b + s_ab;
      }
    double M[D][D];  // G^T G
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = 0; b < D; ++b) {
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) s += G[d][a] * G[d][b];
        M[a][b] = s;
      }
    double t = 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = 0; b < D; ++b) t += Sh[a][b] * M[b][a];
    // A = G Sh M
    double GS[D][D];
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
      for (int b = 0; b < D; ++b) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < D; ++a) s += G[d][a] * Sh[a][b];
        GS[d][b] = s;
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
    const double V = adet;  // W is folded into coef (W, kappa*W or sum_b kappa_b wphi_b)
    dc = (FORM == TM_FORM_DIFFUSION_NODE ? 1.0 : c.W) * V * t;
    const double f = coef * V;
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
      for (int k = 0; k < D; ++k) {
        double A = 0.0;
#pragma unroll
        for (int b = 0; b < D; ++b) A += GS[d][b] * M[b][k];
        dX[d][k] = f * (t * G[d][k] - 2.0 * A);
      }
  }
  if (dcoef) dcoef[e] = (T)dc;
  if (gcoef_slot) {
    // node coefficient: dl/dkappa_b = dc * sum_q w_q phi_b(q)
#pragma unroll
    for (int b = 0; b < B; ++b) gcoef_slot[e * B + b] = (T)(dc * c.wphi[b]);
  }
  if (gx) {
    T* out = gx + e * (B * D);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        out[(k + 1) * D + d] = (T)dX[d][k];
        s += dX[d][k];
      }
      out[d] = (T)(-s);
    }
  }
}

template <typename T, int D>
int launch(const tm_mesh_view* m, const tm_form_desc* f, const int32_t* slot2nnz, const void* gvals, void* dcoef,
           void* gx, void* gcoef_slot, hipStream_t s) {
  constexpr int B = D + 1;
  AdjConsts c{};
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
  const int64_t grid = (m->n_elem + kBlock - 1) / kBlock;
  if (grid == 0) return TM_OK;
  const T* pts = static_cast<const T*>(m->points);
  const T* coeff = static_cast<const T*>(f->coeff);
  const T* gv = static_cast<const T*>(gvals);
#define TM_ADJ(FORM)                                                                                     \
  p1_adjoint_kernel<T, D, FORM><<<grid, kBlock, 0, s>>>(m->conn, pts, coeff, c, slot2nnz, gv, m->n_elem, \
                                                         (T*)dcoef, (T*)gx, (T*)gcoef_slot)
  switch (f->form) {
    case TM_FORM_LAPLACE: TM_ADJ(TM_FORM_LAPLACE); break;
    case TM_FORM_DIFFUSION_ELEM: TM_ADJ(TM_FORM_DIFFUSION_ELEM); break;
    case TM_FORM_DIFFUSION_NODE: TM_ADJ(TM_FORM_DIFFUSION_NODE); break;
    case TM_FORM_MASS: TM_ADJ(TM_FORM_MASS); break;
    default: return tmh::fail(TM_ERR_UNSUPPORTED, "no adjoint for this form");
  }
#undef TM_ADJ
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("adjoint kernel launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
}

}  // namespace

extern "C" int tm_assemble_p1_adjoint(const tm_mesh_view* mesh, const tm_form_desc* form, const int32_t* slot2nnz,
                                      const void* grad_values, const tm_incidence_view* inc, void* ws_points,
                                      void* ws_coeff, void* grad_points, void* grad_coeff, tm_stream_t stream_) {
  TM_TRY_BEGIN
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  if (!mesh || !form || !slot2nnz || !grad_values) return tmh::fail(TM_ERR_ARG, "NULL argument");
  if (mesh->cell != TM_TRI3 && mesh->cell != TM_TET4) return tmh::fail(TM_ERR_UNSUPPORTED, "adjoint is P1 only");
  if (mesh->n_elem * 16 >= tmh::kMaxIndex) return tmh::fail(TM_ERR_SIZE, "mesh too large");
  const int D = tmh::cell_dim(mesh->cell);
  const bool node_coef = form->form == TM_FORM_DIFFUSION_NODE;
  if ((grad_points || node_coef) && !inc) return tmh::fail(TM_ERR_ARG, "incidence needed for node reductions");
  if (grad_points && !ws_points) return tmh::fail(TM_ERR_ARG, "ws_points needed for grad_points");
  if (grad_coeff && node_coef && !ws_coeff) return tmh::fail(TM_ERR_ARG, "ws_coeff needed for a nodal coefficient");
  void* dcoef = (grad_coeff && form->form == TM_FORM_DIFFUSION_ELEM) ? grad_coeff : nullptr;
  void* gslot = (grad_coeff && node_coef) ? ws_coeff : nullptr;
  void* gx = grad_points ? ws_points : nullptr;
  int st;
  if (mesh->dtype == TM_F64)
    st = D == 3 ? launch<double, 3>(mesh, form, slot2nnz, grad_values, dcoef, gx, gslot, stream)
                : launch<double, 2>(mesh, form, slot2nnz, grad_values, dcoef, gx, gslot, stream);
  else
    st = D == 3 ? launch<float, 3>(mesh, form, slot2nnz, grad_values, dcoef, gx, gslot, stream)
                : launch<float, 2>(mesh, form, slot2nnz, grad_values, dcoef, gx, gslot, stream);
  if (st) return st;
  if (gx) {
    st = tm_segmented_reduce(mesh->dtype, mesh->n_points, D, inc->inc_ptr, inc->inc, 0, nullptr, gx, grad_points,
                             stream_);
    if (st) return st;
  }
  if (gslot) {
    st = tm_segmented_reduce(mesh->dtype, mesh->n_points, 1, inc->inc_ptr, inc->inc, inc->n_blocks, inc->blk_ptr,
                             gslot, grad_coeff, stream_);
    if (st) return st;
  }
  return TM_OK;
  TM_TRY_END
}
