// Boundary-facet load vectors (FacetAssembler with a built-in integrand).
//
// Restates FacetAssembler.__call__ (tensormesh/assemble/facet_assembler.py:
// 123-270) for the integrands f * v (f nodal, interpolated at the facet
// points, H values per node) and c * v (c constant, H values): per selected
// facet (element e, local facet f), at each facet quadrature point q, the cell
// Jacobian J[i][j] = sum_b x_b[j] dphi_b/dxi_i (transformation.py:1219-1276)
// at the mapped point, the facet tangents F = (d gamma_f / d r) J and
// FxW = w_q sqrt(det(F F^T)) (transformation.py:985-1097); the facet integral
// I[b][h] = sum_q (f_h(q) phi_b(q)) FxW(q) (the einsum 'fqb...,fq->fb...')
// goes to a slot buffer [facet][b][h], which tm_segmented_reduce sums per
// node over the facet-node incidence lists (ascending facet*B + b: the order
// of the reference's index_add_ over the cell nodes of the selected facets,
// facet_assembler.py:520-525).
#include <hip/hip_runtime.h>

#include "tm_internal.h"

namespace {

constexpr int kBlock = 256;
constexpr int kMaxH = 4;

// table: w [Qf] | phi [nF][Qf][B] | dphi [nF][Qf][D][B] | dgamma [nF][D-1][D]
template <typename T, int D, int B>
__global__ void __launch_bounds__(kBlock)
    facet_load_kernel(const int32_t* __restrict__ conn, const T* __restrict__ pts,
                      const int32_t* __restrict__ felem, const int32_t* __restrict__ flocal, int64_t n_sel,
                      const double* __restrict__ tab, int Qf, int nF, int H, const T* __restrict__ data,
                      const double* __restrict__ cval, T* __restrict__ slot) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n_sel) return;
  const int64_t e = felem[i];
  const int f = flocal[i];
  const double* w = tab;
  const double* phi = tab + Qf + (int64_t)f * Qf * B;
  const double* dphi = tab + Qf + (int64_t)nF * Qf * B + (int64_t)f * Qf * D * B;
  const double* dg = tab + Qf + (int64_t)nF * Qf * B + (int64_t)nF * Qf * D * B + (int64_t)f * (D - 1) * D;
  int v[B];
  double x[B][D];
#pragma unroll
  for (int b = 0; b < B; ++b) {
    v[b] = conn[e * B + b];
#pragma unroll
    for (int d = 0; d < D; ++d) x[b][d] = (double)pts[(int64_t)v[b] * D + d];
  }
  double acc[B][kMaxH];
#pragma unroll
  for (int b = 0; b < B; ++b)
#pragma unroll
    for (int h = 0; h < kMaxH; ++h) acc[b][h] = 0.0;
  for (int q = 0; q < Qf; ++q) {
    const double* ph = phi + q * B;
    const double* dp = dphi + q * D * B;
    double J[D][D];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < B; ++b) s += x[b][j] * dp[a * B + b];
        J[a][j] = s;
      }
    double Tn[D - 1][D];
#pragma unroll
    for (int k = 0; k < D - 1; ++k)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < D; ++a) s += dg[k * D + a] * J[a][j];
        Tn[k][j] = s;
      }
    double det;
    if constexpr (D == 2) {
      det = Tn[0][0] * Tn[0][0] + Tn[0][1] * Tn[0][1];
    } else {
      double g00 = 0.0, g01 = 0.0, g11 = 0.0;
#pragma unroll
      for (int j = 0; j < D; ++j) {
        g00 += Tn[0][j] * Tn[0][j];
        g01 += Tn[0][j] * Tn[1][j];
        g11 += Tn[1][j] * Tn[1][j];
      }
      det = g00 * g11 - g01 * g01;
    }
    const double fxw = w[q] * fabs(sqrt(det));
    for (int h = 0; h < H; ++h) {
      double fh;
      if (data) {
        fh = 0.0;
#pragma unroll
        for (int b = 0; b < B; ++b) fh += (double)data[(int64_t)v[b] * H + h] * ph[b];
      } else {
        fh = cval[h];
      }
#pragma unroll
      for (int b = 0; b < B; ++b) acc[b][h] += fh * ph[b] * fxw;
    }
  }
  T* out = slot + i * B * H;
#pragma unroll
  for (int b = 0; b < B; ++b)
    for (int h = 0; h < H; ++h) out[b * H + h] = (T)acc[b][h];
}

template <typename T, int D, int B>
int launch_facet(const tm_mesh_view* m, const tm_facet_view* fv, const double* tab, int Qf, int H, const void* data,
                 const double* cval, void* slot, hipStream_t s) {
  const int nF = tmh::cell_facets(m->cell);
  const int64_t grid = (fv->n_sel + kBlock - 1) / kBlock;
  if (grid == 0) return TM_OK;
  facet_load_kernel<T, D, B><<<grid, kBlock, 0, s>>>(m->conn, static_cast<const T*>(m->points), fv->elem,
                                                    fv->local, fv->n_sel, tab, Qf, nF, H,
                                                    static_cast<const T*>(data), cval, static_cast<T*>(slot));
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("facet kernel launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
}

template <typename T>
int dispatch_facet(const tm_mesh_view* m, const tm_facet_view* fv, const double* tab, int Qf, int H,
                   const void* data, const double* cval, void* slot, hipStream_t s) {
  switch (m->cell) {
#define TM_CASE(C, D, B) \
  case C: return launch_facet<T, D, B>(m, fv, tab, Qf, H, data, cval, slot, s);
    TM_ALL_CELLS(TM_CASE)
#undef TM_CASE
    default: return tmh::fail(TM_ERR_UNSUPPORTED, "cell type");
  }
}

}  // namespace

extern "C" int tm_assemble_facet_linear(const tm_mesh_view* mesh, const tm_facet_view* facets,
                                        const double* facet_tab_dev, int32_t n_quad, int32_t n_comp,
                                        const void* data, const double* cval_dev, const tm_incidence_view* inc,
                                        void* slot_ws, void* out, tm_stream_t stream_) {
  TM_TRY_BEGIN
  hipStream_t s = reinterpret_cast<hipStream_t>(stream_);
  if (!mesh || !facets || !facet_tab_dev || !inc || !out) return tmh::fail(TM_ERR_ARG, "NULL argument");
  if (!data && !cval_dev) return tmh::fail(TM_ERR_ARG, "the integrand needs nodal data or constants");
  if (n_comp < 1 || n_comp > kMaxH) return tmh::fail(TM_ERR_UNSUPPORTED, "1 to 4 components per node");
  if (n_quad < 1 || n_quad > 16) return tmh::fail(TM_ERR_UNSUPPORTED, "facet rule has too many points");
  if (mesh->dtype != TM_F64 && mesh->dtype != TM_F32) return tmh::fail(TM_ERR_ARG, "unknown dtype");
  const int64_t B = tmh::cell_basis(mesh->cell);
  if (B < 0) return tmh::fail(TM_ERR_ARG, "unknown cell");
  if (facets->n_sel < 0 || facets->n_sel * B * n_comp >= tmh::kMaxIndex) return tmh::fail(TM_ERR_SIZE, "too large");
  if (facets->n_sel == 0)  // no facet selected: a zero load
    return tmh::zero_fill(out, mesh->n_points * n_comp * (mesh->dtype == TM_F64 ? 8 : 4), stream_);
  if (!facets->elem || !facets->local || !slot_ws || !mesh->conn || !mesh->points)
    return tmh::fail(TM_ERR_ARG, "NULL buffer");
  int st = mesh->dtype == TM_F64
               ? dispatch_facet<double>(mesh, facets, facet_tab_dev, n_quad, n_comp, data, cval_dev, slot_ws, s)
               : dispatch_facet<float>(mesh, facets, facet_tab_dev, n_quad, n_comp, data, cval_dev, slot_ws, s);
  if (st) return st;
  return tm_segmented_reduce(mesh->dtype, mesh->n_points, n_comp, inc->inc_ptr, inc->inc, inc->n_blocks,
                             inc->blk_ptr, slot_ws, out, stream_);
  TM_TRY_END
}
