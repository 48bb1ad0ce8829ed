// P1 simplex assembly, stage G (element geometry + local matrix), and the
// two-stage bilinear entry point.
//
// Stage G replaces, for linear triangles and tetrahedra, the chain
//   Element.eval_shape_grad      tensormesh/element/element.py:956-968
//   Transformation jxw = w|detJ| tensormesh/element/transformation.py:1265-1272
//   vmap'd form + _integrate     tensormesh/assemble/element_assembler.py:190-209,416-438
// For P1 the reference gradients are constant, so J, J^-1 and |det J| are
// per element: with edge vectors a_k = x_{k+1} - x_0 (the rows of the
// reference's J[i][j] = sum_b x[b][j] dphi_b/dxi_i), the columns of J^-1 are
// the cofactor vectors n_k / det, and grad(phi_{k+1}) = n_k / det,
// grad(phi_0) = -sum_k grad(phi_{k+1}).  The quadrature sum of a constant
// integrand is (sum_q w_q) * integrand, so
//   laplace / diffusion:  K_ij = c_e * W * |det| * grad_i . grad_j
//   mass:                 M_ij = |det| * sum_q w_q phi_i(q) phi_j(q)
// where c_e is kappa[e], 1, or sum_b kappa[n_b] sum_q w_q phi_b(q) for a
// nodal coefficient.  The local matrix is written to the slot buffer at
// s = e*B*B + i*B + j (the reference's slot numbering,
// element_assembler.py:811-819).
//
// Stage S (ReduceProjector.__call__, projector.py:95-123) lives in
// tm_reduce.hip.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "tm_device.h"
#include "tm_internal.h"

namespace {

constexpr int kBlock = 256;

struct P1Consts {
  double W;           // sum_q w_q
  double wphi[4];     // sum_q w_q phi_b(q)
  double mref[16];    // sum_q w_q phi_i(q) phi_j(q), row-major B x B
  int32_t redo;       // float32 geometry: redo slivers in float64 (not TM_FLAG_F32_NO_REDO)
};

template <int D>
struct Simplex {
  static constexpr int B = D + 1;
};

// Gather the B vertex coordinates of element e as values of the geometry
// type G (double, or float for float meshes).
template <typename T, int D, typename G>
__device__ __forceinline__ void gather(const int32_t* __restrict__ conn, const T* __restrict__ pts,
                                       int64_t e, int (&v)[D + 1], G (&x)[D + 1][D]) {
  constexpr int B = D + 1;
  if constexpr (D == 3) {
    const int4 c = reinterpret_cast<const int4*>(conn)[e];
    v[0] = c.x; v[1] = c.y; v[2] = c.z; v[3] = c.w;
  } else {
#pragma unroll
    for (int b = 0; b < B; ++b) v[b] = conn[e * B + b];
  }
#pragma unroll
  for (int b = 0; b < B; ++b) {
#pragma unroll
    for (int d = 0; d < D; ++d) x[b][d] = (G)pts[(int64_t)v[b] * D + d];
  }
}

// Edge vectors e[k] = x[k+1] - x[0], cofactor vectors n[k] (columns of
// adj(J)) and det J for the P1 simplex.
template <int D, typename G>
__device__ __forceinline__ G cofactors(const G (&x)[D + 1][D], G (&e)[D][D], G (&n)[D][D]) {
#pragma unroll
  for (int k = 0; k < D; ++k)
#pragma unroll
    for (int d = 0; d < D; ++d) e[k][d] = x[k + 1][d] - x[0][d];
  if constexpr (D == 3) {
    const G(&a)[3] = e[0];
    const G(&b)[3] = e[1];
    const G(&c)[3] = e[2];
    // columns of adj(J) for J with rows a, b, c
    n[0][0] = b[1] * c[2] - b[2] * c[1];
    n[0][1] = b[2] * c[0] - b[0] * c[2];
    n[0][2] = b[0] * c[1] - b[1] * c[0];
    n[1][0] = c[1] * a[2] - c[2] * a[1];
    n[1][1] = c[2] * a[0] - c[0] * a[2];
    n[1][2] = c[0] * a[1] - c[1] * a[0];
    n[2][0] = a[1] * b[2] - a[2] * b[1];
    n[2][1] = a[2] * b[0] - a[0] * b[2];
    n[2][2] = a[0] * b[1] - a[1] * b[0];
    return a[0] * n[0][0] + a[1] * n[0][1] + a[2] * n[0][2];
  } else {
    const G a0 = e[0][0], a1 = e[0][1];
    const G b0 = e[1][0], b1 = e[1][1];
    n[0][0] = b1;  n[0][1] = -b0;
    n[1][0] = -a1; n[1][1] = a0;
    return a0 * b1 - a1 * b0;
  }
}

// Local matrix of one element in registers, B*B entries (returned as
// double).  G is the geometry type: float meshes evaluate in float and redo
// elements that tmh::sliver_f32 flags in double.
template <typename T, int D, int FORM, typename G = typename std::conditional<sizeof(T) == 4, float, double>::type>
__device__ __forceinline__ void p1_local(const int32_t* __restrict__ conn, const T* __restrict__ pts,
                                         const T* __restrict__ coeff, const P1Consts& c, int64_t e,
                                         double (&K)[(D + 1) * (D + 1)]) {
  constexpr int B = D + 1;
  int v[B];
  G x[B][D];
  gather<T, D, G>(conn, pts, e, v, x);
  G ed[D][D], n[D][D];
  const G det = cofactors<D, G>(x, ed, n);
  if constexpr (sizeof(G) == 4) {
    if (c.redo && tmh::sliver_f32<D>(ed, det)) {
      p1_local<T, D, FORM, double>(conn, pts, coeff, c, e, K);
      return;
    }
  }
  const G adet = det < G(0) ? -det : det;
  if constexpr (FORM == TM_FORM_MASS) {
#pragma unroll
    for (int k = 0; k < B * B; ++k) K[k] = (double)(adet * (G)c.mref[k]);
    return;
  } else {
    // grad(phi) * det: g[0] = -sum n[k], g[k+1] = n[k]
    G g[B][D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      G s = G(0);
#pragma unroll
      for (int k = 0; k < D; ++k) {
        g[k + 1][d] = n[k][d];
        s += n[k][d];
      }
      g[0][d] = -s;
    }
    G coef = (G)c.W;
    if constexpr (FORM == TM_FORM_DIFFUSION_ELEM) {
      coef *= (G)coeff[e];
    } else if constexpr (FORM == TM_FORM_DIFFUSION_NODE) {
      G kq = G(0);
#pragma unroll
      for (int b = 0; b < B; ++b) kq += (G)coeff[v[b]] * (G)c.wphi[b];
      coef = kq;
    }
    // c * W * |det| * (n_i/det).(n_j/det) = c * W * (n_i . n_j) / |det|
    const G scale = coef / adet;
#pragma unroll
    for (int i = 0; i < B; ++i) {
#pragma unroll
      for (int j = i; j < B; ++j) {
        G s = G(0);
#pragma unroll
        for (int d = 0; d < D; ++d) s += g[i][d] * g[j][d];
        K[i * B + j] = (double)(scale * s);
        K[j * B + i] = (double)(scale * s);
      }
    }
  }
}

// One thread per element; the block's 256 local matrices are staged in LDS
// and written out as contiguous 16-byte vectors (a per-thread store of B*B
// values would hit B*B different cache lines per wave instruction).
template <typename T, int D, int FORM>
__global__ void __launch_bounds__(kBlock) p1_element_kernel(const int32_t* __restrict__ conn,
                                                            const T* __restrict__ pts,
                                                            const T* __restrict__ coeff, P1Consts c,
                                                            int64_t n_elem, T* __restrict__ slot) {
  constexpr int B = D + 1;
  constexpr int BB = B * B;
  __shared__ T tile[kBlock * BB];
  const int64_t e0 = (int64_t)blockIdx.x * kBlock;
  const int64_t e = e0 + threadIdx.x;
  if (e < n_elem) {
    double K[BB];
    p1_local<T, D, FORM>(conn, pts, coeff, c, e, K);
#pragma unroll
    for (int k = 0; k < BB; ++k) tile[threadIdx.x * BB + k] = (T)K[k];
  }
  __syncthreads();
  const int64_t nvalid = (n_elem - e0) < kBlock ? (n_elem - e0) : kBlock;
  const int64_t nbytes = nvalid * BB * (int64_t)sizeof(T);
  const int64_t nvec = nbytes / 16;
  const uint4* src = reinterpret_cast<const uint4*>(tile);
  uint4* dst = reinterpret_cast<uint4*>(slot + e0 * BB);
  for (int64_t t = threadIdx.x; t < nvec; t += kBlock) dst[t] = src[t];
  const int64_t tail = nbytes - nvec * 16;  // multiple of sizeof(T) < 16
  if (threadIdx.x < tail / (int64_t)sizeof(T)) {
    const int64_t off = nvec * 16 / (int64_t)sizeof(T) + threadIdx.x;
    slot[e0 * BB + off] = tile[off];
  }
}

// Stage G writing in incidence order: the local row K_e[i][0..B) of element e
// goes to slot_inc[inc_inv[e*B + i] * B], the position of incidence (e, i) in
// the node-sorted incidence list, so that the row-centric stage S reads every
// CSR row's contributions as one contiguous run.
template <typename T, int D, int FORM>
__global__ void __launch_bounds__(kBlock) p1_element_inc_kernel(const int32_t* __restrict__ conn,
                                                                const T* __restrict__ pts,
                                                                const T* __restrict__ coeff, P1Consts c,
                                                                const int32_t* __restrict__ inc_inv,
                                                                int64_t n_elem, T* __restrict__ slot_inc) {
  constexpr int B = D + 1;
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= n_elem) return;
  double K[B * B];
  p1_local<T, D, FORM>(conn, pts, coeff, c, e, K);
  int32_t dst[B];
  if constexpr (B == 4) {
    const int4 q = reinterpret_cast<const int4*>(inc_inv)[e];
    dst[0] = q.x; dst[1] = q.y; dst[2] = q.z; dst[3] = q.w;
  } else {
#pragma unroll
    for (int i = 0; i < B; ++i) dst[i] = inc_inv[e * B + i];
  }
#pragma unroll
  for (int i = 0; i < B; ++i) {
    T* out = slot_inc + (int64_t)dst[i] * B;
    if constexpr (B == 4 && sizeof(T) == 8) {
      reinterpret_cast<double2*>(out)[0] = make_double2(K[i * B + 0], K[i * B + 1]);
      reinterpret_cast<double2*>(out)[1] = make_double2(K[i * B + 2], K[i * B + 3]);
    } else if constexpr (B == 4 && sizeof(T) == 4) {
      reinterpret_cast<float4*>(out)[0] =
          make_float4((float)K[i * B + 0], (float)K[i * B + 1], (float)K[i * B + 2], (float)K[i * B + 3]);
    } else {
#pragma unroll
      for (int j = 0; j < B; ++j) out[j] = (T)K[i * B + j];
    }
  }
}

int fill_consts(const tm_form_desc* f, int B, int D, P1Consts& c) {
  c = P1Consts{};
  if (!f->tab_host || f->n_quad <= 0) return tmh::fail(TM_ERR_ARG, "form tables missing");
  const int Q = f->n_quad;
  const double* w = f->tab_host;
  const double* phi = f->tab_host + Q;
  for (int q = 0; q < Q; ++q) {
    c.W += w[q];
    for (int b = 0; b < B; ++b) c.wphi[b] += w[q] * phi[q * B + b];
    for (int i = 0; i < B; ++i)
      for (int j = 0; j < B; ++j) c.mref[i * B + j] += w[q] * phi[q * B + i] * phi[q * B + j];
  }
  c.redo = (f->flags & TM_FLAG_F32_NO_REDO) ? 0 : 1;
  (void)D;
  return TM_OK;
}

template <typename T, int D>
int launch_p1(const tm_mesh_view* m, const tm_form_desc* f, void* slot_ws, hipStream_t stream) {
  constexpr int B = D + 1;
  P1Consts c;
  int st = fill_consts(f, B, D, c);
  if (st) return st;
  const int64_t grid = (m->n_elem + kBlock - 1) / kBlock;
  if (grid == 0) return TM_OK;
  const T* pts = static_cast<const T*>(m->points);
  const T* coeff = static_cast<const T*>(f->coeff);
  T* slot = static_cast<T*>(slot_ws);
  switch (f->form) {
    case TM_FORM_LAPLACE:
      p1_element_kernel<T, D, TM_FORM_LAPLACE><<<grid, kBlock, 0, stream>>>(m->conn, pts, coeff, c, m->n_elem, slot);
      break;
    case TM_FORM_DIFFUSION_ELEM:
      if (!coeff) return tmh::fail(TM_ERR_ARG, "diffusion form needs a coefficient");
      p1_element_kernel<T, D, TM_FORM_DIFFUSION_ELEM><<<grid, kBlock, 0, stream>>>(m->conn, pts, coeff, c, m->n_elem, slot);
      break;
    case TM_FORM_DIFFUSION_NODE:
      if (!coeff) return tmh::fail(TM_ERR_ARG, "diffusion form needs a coefficient");
      p1_element_kernel<T, D, TM_FORM_DIFFUSION_NODE><<<grid, kBlock, 0, stream>>>(m->conn, pts, coeff, c, m->n_elem, slot);
      break;
    case TM_FORM_MASS:
      p1_element_kernel<T, D, TM_FORM_MASS><<<grid, kBlock, 0, stream>>>(m->conn, pts, coeff, c, m->n_elem, slot);
      break;
    default:
      return tmh::fail(TM_ERR_UNSUPPORTED, "form not available for P1 scalar assembly");
  }
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("p1 element kernel launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
}

int check_mesh(const tm_mesh_view* m) {
  if (!m) return tmh::fail(TM_ERR_ARG, "mesh is NULL");
  if (tmh::cell_basis(m->cell) < 0) return tmh::fail(TM_ERR_ARG, "unknown cell type");
  if (m->dtype != TM_F32 && m->dtype != TM_F64) return tmh::fail(TM_ERR_ARG, "unknown dtype");
  if (m->n_elem < 0 || m->n_points < 0) return tmh::fail(TM_ERR_ARG, "negative size");
  const int64_t B = tmh::cell_basis(m->cell);
  if (m->n_elem * B * B * 9 >= tmh::kMaxIndex * 4 || m->n_points >= tmh::kMaxIndex)
    return tmh::fail(TM_ERR_SIZE, "mesh too large for int32 indexing");
  if (m->n_elem > 0 && (!m->conn || !m->points)) return tmh::fail(TM_ERR_ARG, "NULL conn / points");
  return TM_OK;
}

}  // namespace

// Defined in tm_p2.hip (P2 and block forms).
int tm_p2_element_matrices(const tm_mesh_view* mesh, const tm_form_desc* form, void* slot_ws,
                           hipStream_t stream, int* hh_out, const int32_t* slot_dest);

namespace {
// element matrices into slot_ws; slot_dest (symmetric blocks only): their
// contribution-ordered positions
int element_matrices(const tm_mesh_view* mesh, const tm_form_desc* form, void* slot_ws, tm_stream_t stream_,
                     const int32_t* slot_dest) {
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  int st = check_mesh(mesh);
  if (st) return st;
  if (!form) return tmh::fail(TM_ERR_ARG, "form is NULL");
  if (mesh->n_elem > 0 && !slot_ws) return tmh::fail(TM_ERR_ARG, "slot workspace is NULL");
  const bool p1_scalar = (mesh->cell == TM_TRI3 || mesh->cell == TM_TET4) && form->form != TM_FORM_LINEAR_ELASTICITY;
  if (p1_scalar && (form->flags & TM_FLAG_UPPER))
    return tmh::fail(TM_ERR_ARG, "TM_FLAG_UPPER: the P1 scalar kernels store every element block");
  if (!p1_scalar) {
    int hh = 0;
    return tm_p2_element_matrices(mesh, form, slot_ws, stream, &hh, slot_dest);
  }
  if (mesh->dtype == TM_F64) {
    return mesh->cell == TM_TET4 ? launch_p1<double, 3>(mesh, form, slot_ws, stream)
                                 : launch_p1<double, 2>(mesh, form, slot_ws, stream);
  }
  return mesh->cell == TM_TET4 ? launch_p1<float, 3>(mesh, form, slot_ws, stream)
                               : launch_p1<float, 2>(mesh, form, slot_ws, stream);
}
}  // namespace

extern "C" int tm_element_matrices(const tm_mesh_view* mesh, const tm_form_desc* form, void* slot_ws,
                                   tm_stream_t stream_) {
  TM_TRY_BEGIN
  return element_matrices(mesh, form, slot_ws, stream_, nullptr);
  TM_TRY_END
}

extern "C" int tm_assemble_bilinear(const tm_mesh_view* mesh, const tm_form_desc* form,
                                    const tm_pattern_view* pat, void* slot_ws, void* values,
                                    tm_stream_t stream) {
  TM_TRY_BEGIN
  if (!pat) return tmh::fail(TM_ERR_ARG, "pattern is NULL");
  if (pat->slot_dest && !(form && (form->flags & TM_FLAG_UPPER)))
    return tmh::fail(TM_ERR_ARG, "slot_dest needs TM_FLAG_UPPER (symmetric element blocks)");
  int st = element_matrices(mesh, form, slot_ws, stream, pat->slot_dest);
  if (st) return st;
  const int hh = form->form == TM_FORM_LINEAR_ELASTICITY ? tmh::cell_dim(mesh->cell) * tmh::cell_dim(mesh->cell) : 1;
  return tmh::reduce_view(mesh->dtype, pat, hh, slot_ws, values, stream);
  TM_TRY_END
}

// Defined in tm_rows.hip: row-centric stage S over an incidence-ordered slot buffer.
int tm_rows_reduce_launch(int32_t dtype, int32_t cell, const tm_rows_view* rows, const void* slot, bool inc_order,
                          void* values, hipStream_t stream);

extern "C" int tm_element_matrices_inc(const tm_mesh_view* mesh, const tm_form_desc* form, const int32_t* inc_inv,
                                       void* slot_ws, tm_stream_t stream_) {
  TM_TRY_BEGIN
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  int st = check_mesh(mesh);
  if (st) return st;
  if (!form || !inc_inv) return tmh::fail(TM_ERR_ARG, "NULL argument");
  if (mesh->cell != TM_TRI3 && mesh->cell != TM_TET4) return tmh::fail(TM_ERR_UNSUPPORTED, "incidence path is P1 only");
  if (mesh->n_elem > 0 && !slot_ws) return tmh::fail(TM_ERR_ARG, "NULL buffer");
  const int B = tmh::cell_basis(mesh->cell), D = tmh::cell_dim(mesh->cell);
  P1Consts c;
  st = fill_consts(form, B, D, c);
  if (st) return st;
  const int64_t grid = (mesh->n_elem + kBlock - 1) / kBlock;
  if (grid > 0) {
#define TM_INC(T, DD, FORM)                                                                                    \
  p1_element_inc_kernel<T, DD, FORM><<<grid, kBlock, 0, stream>>>(mesh->conn, (const T*)mesh->points,         \
                                                                   (const T*)form->coeff, c, inc_inv, mesh->n_elem, \
                                                                   (T*)slot_ws)
#define TM_INC_FORMS(T, DD)                                                                  \
  switch (form->form) {                                                                      \
    case TM_FORM_LAPLACE: TM_INC(T, DD, TM_FORM_LAPLACE); break;                             \
    case TM_FORM_DIFFUSION_ELEM: TM_INC(T, DD, TM_FORM_DIFFUSION_ELEM); break;               \
    case TM_FORM_DIFFUSION_NODE: TM_INC(T, DD, TM_FORM_DIFFUSION_NODE); break;               \
    case TM_FORM_MASS: TM_INC(T, DD, TM_FORM_MASS); break;                                   \
    default: return tmh::fail(TM_ERR_UNSUPPORTED, "form not available for the incidence path"); \
  }
    if ((form->form == TM_FORM_DIFFUSION_ELEM || form->form == TM_FORM_DIFFUSION_NODE) && !form->coeff)
      return tmh::fail(TM_ERR_ARG, "diffusion form needs a coefficient");
    if (mesh->dtype == TM_F64) {
      if (D == 3) { TM_INC_FORMS(double, 3) } else { TM_INC_FORMS(double, 2) }
    } else {
      if (D == 3) { TM_INC_FORMS(float, 3) } else { TM_INC_FORMS(float, 2) }
    }
#undef TM_INC_FORMS
#undef TM_INC
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
      tmh::set_error("incidence element kernel launch: %s", hipGetErrorString(err));
      return TM_ERR_HIP;
    }
  }
  return TM_OK;
  TM_TRY_END
}

extern "C" int tm_assemble_p1_incidence(const tm_mesh_view* mesh, const tm_form_desc* form, const tm_rows_view* rows,
                                        const int32_t* inc_inv, void* slot_ws, void* values, tm_stream_t stream_) {
  TM_TRY_BEGIN
  if (!rows || !values) return tmh::fail(TM_ERR_ARG, "NULL argument");
  int st = tm_element_matrices_inc(mesh, form, inc_inv, slot_ws, stream_);
  if (st) return st;
  return tm_rows_reduce_launch(mesh->dtype, mesh->cell, rows, slot_ws, true, values,
                               reinterpret_cast<hipStream_t>(stream_));
  TM_TRY_END
}
