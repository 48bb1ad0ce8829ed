// J2 (von Mises) plasticity with linear isotropic hardening: incremental
// potential, its gradient (residual), its consistent Hessian (tangent) and
// the radial-return history update, on one wavefront per element.
//
// Restates J2Plasticity (tensormesh/assemble/builtin.py:335-561) evaluated
// through ElementAssembler.energy (element_assembler.py:447-659):
//   G = grad u at each quadrature point (row = component, column = derivative),
//   eps = sym(G) padded to 3 x 3 in 2D, eps_tr = eps - eps_p_n,
//   e = dev(eps_tr), |e| its Frobenius norm,
//   f = 2 mu |e| - sqrt(2/3) (sig0 + H alpha_n), dg = max(f, 0) / (2 mu + 2H/3),
//   psi = K/2 tr(eps_tr)^2 + mu |e|^2 - (2 mu + 2H/3) dg^2 / 2,
// energy = sum_e sum_q psi jxw (the reference's sum over [E, Q]).
//
// The reference obtains the residual by autograd of the energy; here the
// derivatives are written out.  dpsi/dG = K tr delta + c1 sym(e) with
// c1 = 2 mu (1 - dg / |e|) past yield (f >= 0: torch.clamp passes the
// gradient at 0) and c1 = 2 mu below it.  Differentiating once more, with
// m = sym(e) / |e| and c2 = 2 mu (2 mu / (2 mu + 2H/3) - dg / |e|) past yield
// (0 below), the element tangent block (i, j) is
//   sum_q jxw [ (K - c1/3) g_i[a] g_j[b] + c1/2 g_i[b] g_j[a]
//               + c1/2 delta_ab g_i.g_j - c2 (m g_i)[a] (m g_j)[b] ],
// which is linear elasticity (lam, mu) below yield.  At |e| = 0 the
// kernels use that elastic limit; torch autograd's Hessian there is 0 or NaN
// (the derivative of torch.norm at 0).
//
// Wave layout (as the Neo-Hookean kernel): lanes load the element into LDS;
// lane q forms J_q, its inverse and jxw_q; lanes over (q, b, d) form the
// physical gradients; lane q evaluates the material point (and, in update
// mode, writes the new history); lanes over (i, a) sum the residual row and
// lanes over (i, j) the D x D tangent blocks, each stored from its lane's
// registers (with TM_FLAG_UPPER only the pairs i <= j).
#include <hip/hip_runtime.h>

#include "tm_device.h"
#include "tm_internal.h"

namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
static_assert(tmh::kSumParts == TM_J2_ENERGY_PARTS, "energy workspace size");

struct J2Consts {
  double mu, bulk, sig0, hard, denom, r23;
};

// EPW elements per wavefront, each on L = 64 / EPW lanes (the small cells,
// whose stages use at most 16 lanes, pack four with a tangent, eight without).
// TAN: the launch assembles the tangent (the h = m g arrays exist only then).
// BIG (tetra20, hexahedron27): one wavefront per workgroup, rule tables read
// through the caches instead of LDS.
template <typename T, int D, int B, int MQ, int EPW, bool TAN, bool BIG>
__global__ void __launch_bounds__(kWave * (BIG ? 1 : kWavesPerBlock))
    j2_wave_kernel(const int32_t* __restrict__ conn, const T* __restrict__ pts, const T* __restrict__ disp,
                   const double* __restrict__ tab, int Q, J2Consts k, int64_t n_elem, T* __restrict__ eps_p,
                   T* __restrict__ alpha, int update, int upper, double* __restrict__ e_out,
                   T* __restrict__ rslot, T* __restrict__ tslot, const int32_t* __restrict__ slot_dest) {
  constexpr int kWpb = BIG ? 1 : kWavesPerBlock;  // waves per workgroup
  static_assert(!BIG || EPW == 1, "one element per wavefront");
  __shared__ double s_tab[BIG ? 1 : MQ * (1 + B + D * B)];
  __shared__ double s_x[kWpb * EPW][B * D];
  __shared__ double s_u[kWpb * EPW][B * D];
  __shared__ double s_inv[kWpb * EPW][MQ][D * D];
  __shared__ double s_jxw[kWpb * EPW][MQ];
  __shared__ double s_g[kWpb * EPW][MQ * B * D];
  __shared__ double s_h[kWpb * EPW][TAN ? MQ * B * D : 1];
  __shared__ double s_P[kWpb * EPW][MQ][D * D];
  __shared__ double s_m[kWpb * EPW][MQ][D * D];
  __shared__ double s_sc[kWpb * EPW][MQ][4];  // K - c1/3, c1/2, c2, psi jxw
  __shared__ int s_v[kWpb * EPW][B];
  __shared__ uint16_t s_pair[TAN ? B * B : 1];

  constexpr int L = kWave / EPW;
  static_assert(B <= L && MQ <= L, "an element's lanes index its nodes and its quadrature points");
  const int lane = threadIdx.x % L;   // lane within the element's group
  const int wv = threadIdx.x / L;     // element slot within the workgroup
  const int64_t e = (int64_t)blockIdx.x * (kWpb * EPW) + wv;
  const bool active = e < n_elem;
  const int tab_len = Q + Q * B + Q * D * B;
  if constexpr (!BIG)
    for (int t = threadIdx.x; t < tab_len; t += blockDim.x) s_tab[t] = tab[t];
  int n_pairs = 0;
  if constexpr (TAN) n_pairs = tmh::fill_pairs<B>(s_pair, upper != 0);
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
    dst[r] = (TAN && !update && slot_dest && active && p < n_pairs) ? slot_dest[e * n_pairs + p] : 0;
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
  if (active && lane < Q) {
    const int q = lane;
    const double* g = &s_g[wv][q * B * D];
    double G[D][D];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int d = 0; d < D; ++d) {
        double s = 0.0;
        for (int b = 0; b < B; ++b) s += s_u[wv][b * D + a] * g[b * D + d];
        G[a][d] = s;
      }
    const int64_t pq = e * Q + q;
    T* ep = eps_p + pq * 9;
    double x[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double es = (r < D && c < D) ? 0.5 * (G[r][c] + G[c][r]) : 0.0;
        x[r][c] = es - (double)ep[r * 3 + c];
      }
    const double tr = x[0][0] + x[1][1] + x[2][2];
    double nrm2 = 0.0;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      x[r][r] -= tr / 3.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) nrm2 += x[r][c] * x[r][c];
    }
    const double nrm = sqrt(nrm2);
    const double al = (double)alpha[pq];
    const double f = 2.0 * k.mu * nrm - k.r23 * (k.sig0 + k.hard * al);
    const double dg = fmax(f, 0.0) / k.denom;
    if (update) {
      // builtin.py:513-561: eps_p += dg n, alpha += sqrt(2/3) dg where f > 0,
      // n = e / |e| (|e| < 1e-12 divides by 1)
      if (f > 0.0) {
        const double inv_n = 1.0 / (nrm < 1e-12 ? 1.0 : nrm);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) ep[r * 3 + c] = (T)((double)ep[r * 3 + c] + dg * (x[r][c] * inv_n));
        alpha[pq] = (T)(al + k.r23 * dg);
      }
    } else {
      const double psi = 0.5 * k.bulk * tr * tr + k.mu * nrm2 - 0.5 * k.denom * dg * dg;
      const bool plastic = f >= 0.0 && nrm > 0.0;
      const double c1 = plastic ? 2.0 * k.mu * (1.0 - dg / nrm) : 2.0 * k.mu;
      const double c2 = plastic ? 2.0 * k.mu * (2.0 * k.mu / k.denom - dg / nrm) : 0.0;
      const double inv_n = plastic ? 1.0 / nrm : 0.0;
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int A = 0; A < D; ++A) {
          const double se = 0.5 * (x[a][A] + x[A][a]);
          s_P[wv][q][a * D + A] = (a == A ? k.bulk * tr : 0.0) + c1 * se;
          s_m[wv][q][a * D + A] = se * inv_n;
        }
      s_sc[wv][q][0] = k.bulk - c1 / 3.0;
      s_sc[wv][q][1] = 0.5 * c1;
      s_sc[wv][q][2] = c2;
      s_sc[wv][q][3] = psi * s_jxw[wv][q];
    }
  }
  __syncthreads();
  if (!update && active) {
    if (e_out && lane == 0) {
      double s = 0.0;
      for (int q = 0; q < Q; ++q) s += s_sc[wv][q][3];
      e_out[e] = s;
    }
    if (rslot) {
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
        rslot[(e * B + i) * D + a] = (T)acc;
      }
    }
    if (TAN) {
      for (int t = lane; t < Q * B * D; t += L) {
        const int q = t / (B * D), b = (t / D) % B, a = t % D;
        const double* g = &s_g[wv][(q * B + b) * D];
        double s = 0.0;
#pragma unroll
        for (int A = 0; A < D; ++A) s += s_m[wv][q][a * D + A] * g[A];
        s_h[wv][t] = s;
      }
    }
  }
  __syncthreads();
  if (TAN && !update && active) {
    constexpr int DD = D * D;
#pragma unroll
    for (int r = 0; r < kPairRounds; ++r) {
      const int p = lane + r * L;
      if (p >= n_pairs) break;
      int i = s_pair[p] >> 8, j = s_pair[p] & 255;
      // block p of the element's pairs (packed when upper): 72 contiguous bytes
      // at f64; with slot_dest at its contribution-ordered slot, transposed
      // when the id is negative -- formed as block (j, i), which is the exact
      // transpose of (i, j) (the same products, commuted)
      T* out = tslot + (e * n_pairs + p) * DD;
      if (slot_dest) {
        const int32_t d = dst[r];
        out = tslot + (int64_t)(d & INT32_MAX) * DD;
        if (d < 0) {
          const int t = i;
          i = j;
          j = t;
        }
      }
      double acc[DD];
#pragma unroll
      for (int c = 0; c < DD; ++c) acc[c] = 0.0;
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
        const double wq = s_jxw[wv][q];
        const double lw = s_sc[wv][q][0] * wq, mw = s_sc[wv][q][1] * wq, cw = s_sc[wv][q][2] * wq;
        double dot = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) dot += vgi[d] * vgj[d];
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
          for (int b = 0; b < D; ++b) {
            double t = lw * (vgi[a] * vgj[b]) + mw * (vgi[b] * vgj[a]) - cw * (vhi[a] * vhj[b]);
            if (a == b) t += mw * dot;
            acc[a * D + b] += t;
          }
      }
#pragma unroll
      for (int c = 0; c < DD; ++c) out[c] = (T)acc[c];
    }
  }
}

template <typename T, int D, int B>
int launch_j2(const tm_mesh_view* m, const tm_form_desc* f, const J2Consts& k, const void* disp,
              const tm_j2_state* st, int update, double* e_out, void* rslot, void* tslot, hipStream_t s,
              const int32_t* slot_dest) {
  // more than 32 element dofs (tetra20: 60, hexahedron27: 81), or a rule of
  // more than 16 points (the 24-point tetrahedron rule): one element per
  // workgroup, the BIG variant, with rules of up to 32 points
  constexpr bool kBigCell = B * D > 32;
  const int Q = f->n_quad;
  const bool big = kBigCell || Q > 16;
  const int wpb = big ? 1 : kWavesPerBlock;
  if (Q <= 0 || Q > 32) return tmh::fail(TM_ERR_UNSUPPORTED, "quadrature rule has too many points");
  if (!f->tab_dev) return tmh::fail(TM_ERR_ARG, "device form tables missing");
  // four small elements (B <= 4) per wavefront when the rule has at most 4 points
  // small cells (B <= 4, rules of <= 4 points): eight elements per wavefront
  // without a tangent, four with one (its pair loop wants the lanes)
  const bool tan = tslot && !update;
  int epw = (B <= 4 && Q <= 4) ? (tan ? 4 : 8) : 1;
  if (f->epw != 0) {
    // form->epw: 1 everywhere; 4 or 8 for cells of <= 4 nodes with rules of <= 4 points
    if (!(f->epw == 1 || ((f->epw == 4 || f->epw == 8) && B <= 4 && Q <= 4)))
      return tmh::fail(TM_ERR_ARG, "elements per wavefront not admitted for this cell / rule");
    epw = f->epw;
  }
  const int64_t per_block = (int64_t)wpb * epw;
  const int64_t grid = (m->n_elem + per_block - 1) / per_block;
  if (grid == 0) return TM_OK;
  const dim3 g((unsigned)grid), blk(kWave * wpb);
  const int upper = (f->flags & TM_FLAG_UPPER) != 0;
#define TM_J2_TAN(MQ, EPW, TAN, BG)                                                                    \
  j2_wave_kernel<T, D, B, MQ, EPW, TAN, BG><<<g, blk, 0, s>>>(                                         \
      m->conn, (const T*)m->points, (const T*)disp, f->tab_dev, Q, k, m->n_elem, (T*)st->eps_p,       \
      (T*)st->alpha, update, upper, e_out, (T*)rslot, (T*)tslot, upper ? slot_dest : nullptr)
#define TM_J2_V(MQ, EPW, BG)          \
  do {                                \
    if (tslot && !update)             \
      TM_J2_TAN(MQ, EPW, true, BG);   \
    else                              \
      TM_J2_TAN(MQ, EPW, false, BG);  \
  } while (0)
#define TM_J2(MQ, EPW) TM_J2_V(MQ, EPW, false)
  if constexpr (kBigCell) {
    if (Q <= 16) TM_J2_V(16, 1, true); else TM_J2_V(32, 1, true);
  } else if (big) {
    TM_J2_V(32, 1, true);
  } else if constexpr (B <= 4) {
    if (epw == 8) TM_J2(4, 8);
    else if (epw == 4) TM_J2(4, 4);
    else if (Q <= 4) TM_J2(4, 1);
    else TM_J2(16, 1);
  } else {
    if (Q <= 4) TM_J2(4, 1); else TM_J2(16, 1);
  }
#undef TM_J2
#undef TM_J2_V
#undef TM_J2_TAN
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("J2 kernel launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
}

template <typename T>
int dispatch_j2(const tm_mesh_view* m, const tm_form_desc* f, const J2Consts& k, const void* disp,
                const tm_j2_state* st, int update, double* e_out, void* rslot, void* tslot, hipStream_t s,
                const int32_t* slot_dest = nullptr) {
  switch (m->cell) {
#define TM_CASE(C, D, B) \
  case C: return launch_j2<T, D, B>(m, f, k, disp, st, update, e_out, rslot, tslot, s, slot_dest);
    TM_ALL_CELLS(TM_CASE)
#undef TM_CASE
    default: return tmh::fail(TM_ERR_ARG, "unknown cell");
  }
}

int check_common(const tm_mesh_view* m, const tm_form_desc* f, const tm_j2_material* mat, const void* disp,
                 const tm_j2_state* st, J2Consts* k) {
  if (!m || !f || !mat || !disp || !st) return tmh::fail(TM_ERR_ARG, "NULL argument");
  if (m->dtype != TM_F64 && m->dtype != TM_F32) return tmh::fail(TM_ERR_ARG, "unknown dtype");
  const int64_t B = tmh::cell_basis(m->cell);
  if (B < 0) return tmh::fail(TM_ERR_ARG, "unknown cell");
  if (m->n_elem < 0 || m->n_elem * B * B >= tmh::kMaxIndex) return tmh::fail(TM_ERR_SIZE, "mesh too large");
  if (m->n_elem > 0 && (!st->eps_p || !st->alpha || !m->conn || !m->points))
    return tmh::fail(TM_ERR_ARG, "NULL buffer");
  if (!(mat->poisson > -1.0 && mat->poisson < 0.5)) return tmh::fail(TM_ERR_ARG, "Poisson ratio outside (-1, 0.5)");
  k->mu = mat->young / (2.0 * (1.0 + mat->poisson));
  const double lam = mat->young * mat->poisson / ((1.0 + mat->poisson) * (1.0 - 2.0 * mat->poisson));
  k->bulk = lam + 2.0 / 3.0 * k->mu;
  k->sig0 = mat->sig0;
  k->hard = mat->hardening;
  k->denom = 2.0 * k->mu + 2.0 / 3.0 * mat->hardening;
  k->r23 = sqrt(2.0 / 3.0);
  if (!(k->denom > 0.0)) return tmh::fail(TM_ERR_ARG, "2 mu + 2H/3 must be positive");
  return TM_OK;
}

}  // namespace

extern "C" int tm_assemble_j2(const tm_mesh_view* mesh, const tm_form_desc* form, const tm_j2_material* mat,
                              const void* disp, const tm_j2_state* state, void* energy_ws, void* energy,
                              const tm_pattern_view* pat, const tm_incidence_view* inc, void* tslot_ws,
                              void* tangent, void* rslot_ws, void* residual, tm_stream_t stream_) {
  TM_TRY_BEGIN
  hipStream_t s = reinterpret_cast<hipStream_t>(stream_);
  J2Consts k;
  int st = check_common(mesh, form, mat, disp, state, &k);
  if (st) return st;
  const int64_t D = tmh::cell_dim(mesh->cell);
  if (energy && !energy_ws) return tmh::fail(TM_ERR_ARG, "energy needs energy_ws [n_elem + 1024] doubles");
  if (tangent && (!pat || !tslot_ws)) return tmh::fail(TM_ERR_ARG, "tangent needs a pattern and a slot workspace");
  if (residual && (!inc || !rslot_ws)) return tmh::fail(TM_ERR_ARG, "residual needs incidence and a workspace");
  if (!energy && !tangent && !residual) return TM_OK;
  double* e_out = energy ? static_cast<double*>(energy_ws) : nullptr;
  void* rs = residual ? rslot_ws : nullptr;
  void* ts = tangent ? tslot_ws : nullptr;
  if (tangent && pat->slot_dest && !(form->flags & TM_FLAG_UPPER))
    return tmh::fail(TM_ERR_ARG, "slot_dest needs TM_FLAG_UPPER (symmetric element blocks)");
  const int32_t* sd = tangent ? pat->slot_dest : nullptr;
  st = mesh->dtype == TM_F64 ? dispatch_j2<double>(mesh, form, k, disp, state, 0, e_out, rs, ts, s, sd)
                             : dispatch_j2<float>(mesh, form, k, disp, state, 0, e_out, rs, ts, s, sd);
  if (st) return st;
  if (energy) {
    st = tmh::sum_fixed(e_out, mesh->n_elem, e_out + mesh->n_elem, mesh->dtype, energy, stream_);
    if (st) return st;
  }
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

extern "C" int tm_j2_update_state(const tm_mesh_view* mesh, const tm_form_desc* form, const tm_j2_material* mat,
                                  const void* disp, const tm_j2_state* state, tm_stream_t stream_) {
  TM_TRY_BEGIN
  hipStream_t s = reinterpret_cast<hipStream_t>(stream_);
  J2Consts k;
  int st = check_common(mesh, form, mat, disp, state, &k);
  if (st) return st;
  return mesh->dtype == TM_F64 ? dispatch_j2<double>(mesh, form, k, disp, state, 1, nullptr, nullptr, nullptr, s)
                               : dispatch_j2<float>(mesh, form, k, disp, state, 1, nullptr, nullptr, nullptr, s);
  TM_TRY_END
}
