// Penalty contact against a plane on the selected boundary facets:
// energy, gradient and Hessian with respect to the nodal positions.
//
// Restates ContactAssembler.energy (tensormesh/assemble/builtin.py:609-728)
// for the penalty density of its example (builtin.py:588-597), with the
// obstacle a plane n . x = offset:
//   x_q = sum_b phi_b(q) x_b  (cell basis at the facet point, 'sbd,sqb->sqd'),
//   g_q = n . x_q - offset,   psi_q = kappa/2 max(-g_q, 0)^2,
//   energy = sum_facets sum_q psi_q FxW_q,
// with FxW_q = w_q sqrt(det(F F^T)) of the configuration in mesh->points
// (the facet tangents F = (d gamma / d r) J, transformation.py:985-1097),
// while the densities see the positions x.  Derivatives as torch autograd
// gives them for torch.clamp (the gradient passes at g = 0):
//   dE/dx_b   = sum_q FxW_q phi_b kappa min(g_q, 0) n,
//   d2E/dx_b dx_c = (sum_q FxW_q phi_b phi_c kappa [g_q <= 0]) n n^T.
// One thread per selected facet writes its energy, its B x D gradient rows
// and its B x B blocks of D x D; fixed-order reductions then sum them per
// node (the facet incidence lists) and per pattern entry.
#include <hip/hip_runtime.h>

#include "tm_internal.h"

namespace {

constexpr int kBlock = 128;
constexpr int kMaxQ = 16;

struct Plane {
  double kappa, n[3], offset;
};

// table: w [Qf] | phi [nF][Qf][B] | dphi [nF][Qf][D][B] | dgamma [nF][D-1][D]
template <typename T, int D, int B>
__global__ void __launch_bounds__(kBlock)
    contact_penalty_kernel(const int32_t* __restrict__ conn, const T* __restrict__ pts, const T* __restrict__ xpos,
                           const int32_t* __restrict__ felem, const int32_t* __restrict__ flocal, int64_t n_sel,
                           const double* __restrict__ tab, int Qf, int nF, Plane pl, double* __restrict__ e_out,
                           T* __restrict__ rslot, T* __restrict__ tslot) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n_sel) return;
  const int64_t e = felem[i];
  const int f = flocal[i];
  const double* w = tab;
  const double* phi = tab + Qf + (int64_t)f * Qf * B;
  const double* dphi = tab + Qf + (int64_t)nF * Qf * B + (int64_t)f * Qf * D * B;
  const double* dg = tab + Qf + (int64_t)nF * Qf * B + (int64_t)nF * Qf * D * B + (int64_t)f * (D - 1) * D;
  double X[B][D], xs[B][D];
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int64_t v = conn[e * B + b];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      X[b][d] = (double)pts[v * D + d];
      xs[b][d] = (double)xpos[v * D + d];
    }
  }
  double energy = 0.0;
  double gr[B];  // sum_q FxW phi_b kappa min(g, 0)
  double cq[kMaxQ];  // FxW kappa [g <= 0] per point
#pragma unroll
  for (int b = 0; b < B; ++b) gr[b] = 0.0;
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
        for (int b = 0; b < B; ++b) s += X[b][j] * dp[a * B + b];
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
    double gap = -pl.offset;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double xq = 0.0;
#pragma unroll
      for (int b = 0; b < B; ++b) xq += xs[b][d] * ph[b];
      gap += pl.n[d] * xq;
    }
    const bool touch = gap <= 0.0;
    const double pen = touch ? -gap : 0.0;
    energy += 0.5 * pl.kappa * pen * pen * fxw;
    const double gcoef = touch ? fxw * pl.kappa * gap : 0.0;
#pragma unroll
    for (int b = 0; b < B; ++b) gr[b] += gcoef * ph[b];
    if (q < kMaxQ) cq[q] = touch ? fxw * pl.kappa : 0.0;
  }
  if (e_out) e_out[i] = energy;
  if (rslot) {
    T* out = rslot + i * B * D;
#pragma unroll
    for (int b = 0; b < B; ++b)
#pragma unroll
      for (int d = 0; d < D; ++d) out[b * D + d] = (T)(gr[b] * pl.n[d]);
  }
  if (tslot) {
    T* out = tslot + i * B * B * D * D;
    for (int b = 0; b < B; ++b)
      for (int c = 0; c < B; ++c) {
        double s = 0.0;
        for (int q = 0; q < Qf; ++q) s += cq[q] * phi[q * B + b] * phi[q * B + c];
        T* blk = out + (b * B + c) * D * D;
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
          for (int d = 0; d < D; ++d) blk[a * D + d] = (T)(s * pl.n[a] * pl.n[d]);
      }
  }
}

template <typename T, int D, int B>
int launch_contact(const tm_mesh_view* m, const tm_facet_view* fv, const double* tab, int Qf, const Plane& pl,
                   const void* x, double* e_out, void* rslot, void* tslot, hipStream_t s) {
  const int64_t grid = (fv->n_sel + kBlock - 1) / kBlock;
  if (grid == 0) return TM_OK;
  const int nF = tmh::cell_facets(m->cell);
  contact_penalty_kernel<T, D, B><<<grid, kBlock, 0, s>>>(m->conn, static_cast<const T*>(m->points),
                                                         static_cast<const T*>(x), fv->elem, fv->local, fv->n_sel,
                                                         tab, Qf, nF, pl, e_out, static_cast<T*>(rslot),
                                                         static_cast<T*>(tslot));
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("contact kernel launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
}

template <typename T>
int dispatch_contact(const tm_mesh_view* m, const tm_facet_view* fv, const double* tab, int Qf, const Plane& pl,
                     const void* x, double* e_out, void* rslot, void* tslot, hipStream_t s) {
  switch (m->cell) {
#define TM_CASE(C, D, B) \
  case C: return launch_contact<T, D, B>(m, fv, tab, Qf, pl, x, e_out, rslot, tslot, s);
    TM_ALL_CELLS(TM_CASE)
#undef TM_CASE
    default: return tmh::fail(TM_ERR_UNSUPPORTED, "cell type");
  }
}

}  // namespace

extern "C" int tm_contact_penalty(const tm_mesh_view* mesh, const tm_facet_view* facets,
                                  const double* facet_tab_dev, int32_t n_quad, const tm_contact_plane* plane,
                                  const void* x, void* energy_ws, void* energy, const tm_incidence_view* inc,
                                  void* rslot_ws, void* grad, const tm_pattern_view* fpat, void* tslot_ws,
                                  void* hess, tm_stream_t stream_) {
  TM_TRY_BEGIN
  hipStream_t s = reinterpret_cast<hipStream_t>(stream_);
  if (!mesh || !facets || !facet_tab_dev || !plane || !x) return tmh::fail(TM_ERR_ARG, "NULL argument");
  if (mesh->dtype != TM_F64 && mesh->dtype != TM_F32) return tmh::fail(TM_ERR_ARG, "unknown dtype");
  if (n_quad < 1 || n_quad > kMaxQ) return tmh::fail(TM_ERR_UNSUPPORTED, "facet rule has too many points");
  const int64_t B = tmh::cell_basis(mesh->cell), D = tmh::cell_dim(mesh->cell);
  if (B < 0) return tmh::fail(TM_ERR_ARG, "unknown cell");
  if (facets->n_sel < 0 || facets->n_sel * B * B * D * D >= tmh::kMaxIndex) return tmh::fail(TM_ERR_SIZE, "too large");
  if (facets->n_sel > 0 && (!facets->elem || !facets->local || !mesh->conn || !mesh->points))
    return tmh::fail(TM_ERR_ARG, "NULL buffer");
  if (energy && !energy_ws) return tmh::fail(TM_ERR_ARG, "energy needs energy_ws [n_sel + 1024] doubles");
  if (grad && (!inc || !rslot_ws)) return tmh::fail(TM_ERR_ARG, "gradient needs incidence and a workspace");
  if (hess && (!fpat || !tslot_ws)) return tmh::fail(TM_ERR_ARG, "Hessian needs facet-slot segments and a workspace");
  if (!(plane->kappa >= 0.0)) return tmh::fail(TM_ERR_ARG, "penalty must be non-negative");
  if (facets->n_sel == 0) {
    // no facet selected: zero energy, gradient and Hessian
    const int64_t s_val = mesh->dtype == TM_F64 ? 8 : 4;
    int st = energy ? tmh::zero_fill(energy, s_val, stream_) : TM_OK;
    if (!st && grad) st = tmh::zero_fill(grad, mesh->n_points * D * s_val, stream_);
    // a subset view (fpat->entry set) names only touched entries: the caller
    // zeroes hess then, so nothing is written; a full view covers every entry
    if (!st && hess && !fpat->entry) st = tmh::zero_fill(hess, fpat->nnz * D * D * s_val, stream_);
    return st;
  }
  Plane pl{plane->kappa, {plane->normal[0], plane->normal[1], plane->normal[2]}, plane->offset};
  double* e_out = energy ? static_cast<double*>(energy_ws) : nullptr;
  void* rs = grad ? rslot_ws : nullptr;
  void* ts = hess ? tslot_ws : nullptr;
  int st = mesh->dtype == TM_F64 ? dispatch_contact<double>(mesh, facets, facet_tab_dev, n_quad, pl, x, e_out, rs, ts, s)
                                 : dispatch_contact<float>(mesh, facets, facet_tab_dev, n_quad, pl, x, e_out, rs, ts, s);
  if (st) return st;
  if (energy) {
    st = tmh::sum_fixed(e_out, facets->n_sel, e_out + facets->n_sel, mesh->dtype, energy, stream_);
    if (st) return st;
  }
  if (grad) {
    st = tm_segmented_reduce(mesh->dtype, mesh->n_points, (int32_t)D, inc->inc_ptr, inc->inc, inc->n_blocks,
                             inc->blk_ptr, rslot_ws, grad, stream_);
    if (st) return st;
  }
  if (hess) {
    st = tmh::reduce_view(mesh->dtype, fpat, (int32_t)(D * D), tslot_ws, hess, stream_);
    if (st) return st;
  }
  return TM_OK;
  TM_TRY_END
}
