// Fused row-centric assembly for P1 scalar forms (no slot buffer).
//
// Same result as the two-stage path of tm_p1.hip + tm_reduce.hip -- geometry
// (element.py:956-968, transformation.py:1265-1272), form and integration
// (element_assembler.py:190-209,416-438), ReduceProjector (projector.py:95-123)
// -- computed from the CSR rows' side.  For node r, its incidence list
// (e, i) (ascending e, conn[e,i] == r) names every element whose local row i
// lands in CSR row r; colpos packs, per incidence, the position of each
// conn[e,j] inside row r.  A workgroup owns a run of whole rows (a reduction
// plan over inc_ptr):
//   1. one thread per incidence recomputes the element's geometry and its
//      local row K_e[i][0..B) and stages it, with its colpos word, in LDS;
//   2. one thread per CSR entry (r, p) walks row r's incidences in order and
//      adds the staged value whose colpos byte equals p.
// Every entry is therefore summed in ascending element order, as the
// reference's sequential index_add_ does, without atomics; each element is
// evaluated B times (once per vertex row) instead of being written to and
// read back from a B*B slot buffer.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "tm_device.h"
#include "tm_internal.h"

namespace {

constexpr int kBlock = 256;
constexpr int kCapInc = 512;   // incidences staged per workgroup (matches the host plan cap)
constexpr int kMaxRows = 64;   // rows staged per workgroup (the host plan splits longer runs)
constexpr int kMaxEnt = 1024;  // CSR entries staged per workgroup

struct RowConsts {
  double W;        // sum_q w_q
  double wphi[4];  // sum_q w_q phi_b(q)
  double mref[16]; // sum_q w_q phi_i phi_j
  int32_t redo;    // float32 geometry: redo slivers in float64 (not TM_FLAG_F32_NO_REDO)
};

__device__ __forceinline__ int64_t xcd_logical(int64_t b, int64_t nb) {
  const int64_t q = nb / 8, r = nb % 8;
  const int64_t x = b % 8, y = b / 8;
  return x < r ? x * (q + 1) + y : r * (q + 1) + (x - r) * q + y;
}

// Coordinates of node v as D values of the geometry type G (double, or
// float for float meshes): for 3-D double points one 16-byte and one 8-byte
// load (8-byte aligned), for float points one 12-byte load.
template <typename T, int D, typename G>
__device__ __forceinline__ void load_node(const T* __restrict__ pts, int32_t v, G (&x)[D]) {
  const T* p = pts + (int64_t)v * D;
  if constexpr (D == 3 && sizeof(T) == 8) {
    typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));
    const d2u a = *reinterpret_cast<const d2u*>(p);
    x[0] = (G)a.x;
    x[1] = (G)a.y;
    x[2] = (G)p[2];
  } else if constexpr (D == 3) {
    typedef float f3u __attribute__((ext_vector_type(3), aligned(4)));
    const f3u a = *reinterpret_cast<const f3u*>(p);
    x[0] = (G)a.x;
    x[1] = (G)a.y;
    x[2] = (G)a.z;
  } else {
#pragma unroll
    for (int j = 0; j < D; ++j) x[j] = (G)p[j];
  }
}

// Geometry type of the row kernels: float meshes evaluate element geometry
// in float and redo the slivers in double (tmh::sliver_f32).
template <typename T>
using Geom = typename std::conditional<sizeof(T) == 4, float, double>::type;

// Element data a local row needs, gathered in three dependent steps
// (incidence -> connectivity -> coordinates / coefficients) so that a thread
// can issue the loads of several incidences before it computes any of them.
template <int D, typename G>
struct ElemData {
  static constexpr int B = D + 1;
  int v[B];
  G x[B][D];
  G k[B];  // element coefficient in k[0], or the nodal coefficients
};

template <int D, typename G>
__device__ __forceinline__ void load_conn(const int32_t* __restrict__ conn, int64_t e, ElemData<D, G>& d) {
  constexpr int B = D + 1;
  if constexpr (D == 3) {
    const int4 q = reinterpret_cast<const int4*>(conn)[e];
    d.v[0] = q.x; d.v[1] = q.y; d.v[2] = q.z; d.v[3] = q.w;
  } else {
#pragma unroll
    for (int b = 0; b < B; ++b) d.v[b] = conn[e * B + b];
  }
}

template <typename T, int D, int FORM, typename G>
__device__ __forceinline__ void load_geometry(const T* __restrict__ pts, const T* __restrict__ coeff, int64_t e,
                                              ElemData<D, G>& d) {
  constexpr int B = D + 1;
#pragma unroll
  for (int b = 0; b < B; ++b) load_node<T, D, G>(pts, d.v[b], d.x[b]);
  if constexpr (FORM == TM_FORM_DIFFUSION_ELEM) {
    d.k[0] = (G)coeff[e];
  } else if constexpr (FORM == TM_FORM_DIFFUSION_NODE) {
#pragma unroll
    for (int b = 0; b < B; ++b) d.k[b] = (G)coeff[d.v[b]];
  }
}

// Rotated element view for the Laplace / element-coefficient forms: for the
// incidence (e, i), vertex p of the view is vertex (i + p) mod B of e, so the
// row's own node is always view vertex 0 and nothing is indexed by the
// runtime i in registers (selects on the connectivity word only).  The row
// node's coordinates come from LDS (staged per row); the other B - 1
// vertices are gathered.
template <int D, typename G>
struct RotData {
  static constexpr int B = D + 1;
  int32_t c[B];  // rotated connectivity
  G x[B][D];     // rotated coordinates (x[0] filled from LDS)
  G k;           // element coefficient
};

template <typename T, int D, int FORM, typename G>
__device__ __forceinline__ void load_rot(const int32_t* __restrict__ conn, const T* __restrict__ coeff, int64_t e,
                                         int i, RotData<D, G>& d) {
  if constexpr (D == 3) {
    const int4 q = reinterpret_cast<const int4*>(conn)[e];
    d.c[0] = i == 0 ? q.x : i == 1 ? q.y : i == 2 ? q.z : q.w;
    d.c[1] = i == 0 ? q.y : i == 1 ? q.z : i == 2 ? q.w : q.x;
    d.c[2] = i == 0 ? q.z : i == 1 ? q.w : i == 2 ? q.x : q.y;
    d.c[3] = i == 0 ? q.w : i == 1 ? q.x : i == 2 ? q.y : q.z;
  } else {
    const int32_t a0 = conn[e * 3 + 0], a1 = conn[e * 3 + 1], a2 = conn[e * 3 + 2];
    d.c[0] = i == 0 ? a0 : i == 1 ? a1 : a2;
    d.c[1] = i == 0 ? a1 : i == 1 ? a2 : a0;
    d.c[2] = i == 0 ? a2 : i == 1 ? a0 : a1;
  }
  if constexpr (FORM == TM_FORM_DIFFUSION_ELEM) d.k = (G)coeff[e];
}

// Connectivity word(s) of element e: one int4 for tetrahedra, three ints
// for triangles (loads only, so that a thread can issue those of several
// elements before it selects from any of them).  Named fields rather than an
// array: the rotation below selects among them, and an array indexed by the
// runtime i would be placed in scratch memory.
template <int D>
__device__ __forceinline__ int4 load_conn_word(const int32_t* __restrict__ conn, int64_t e) {
  if constexpr (D == 3) {
    return reinterpret_cast<const int4*>(conn)[e];
  } else {
    return make_int4(conn[e * 3 + 0], conn[e * 3 + 1], conn[e * 3 + 2], 0);
  }
}

// The rotated view of incidence (e, i) from its loaded connectivity word.
template <int D, typename G>
__device__ __forceinline__ void rotate_conn(const int4 q, int i, RotData<D, G>& d) {
  if constexpr (D == 3) {
    d.c[0] = i == 0 ? q.x : i == 1 ? q.y : i == 2 ? q.z : q.w;
    d.c[1] = i == 0 ? q.y : i == 1 ? q.z : i == 2 ? q.w : q.x;
    d.c[2] = i == 0 ? q.z : i == 1 ? q.w : i == 2 ? q.x : q.y;
    d.c[3] = i == 0 ? q.w : i == 1 ? q.x : i == 2 ? q.y : q.z;
  } else {
    d.c[0] = i == 0 ? q.x : i == 1 ? q.y : q.z;
    d.c[1] = i == 0 ? q.y : i == 1 ? q.z : q.x;
    d.c[2] = i == 0 ? q.z : i == 1 ? q.x : q.y;
  }
}

template <typename T, int D, typename G>
__device__ __forceinline__ void gather_rot(const T* __restrict__ pts, RotData<D, G>& d) {
#pragma unroll
  for (int p = 1; p < D + 1; ++p) load_node<T, D, G>(pts, d.c[p], d.x[p]);
}

// Edge vectors from vertex 0, cofactor vectors n[k] (columns of adj J) and
// det J of the simplex with vertices x.
template <int D, typename G>
__device__ __forceinline__ G simplex_cofactors(const G (&x)[D + 1][D], G (&e)[D][D], G (&n)[D][D]) {
#pragma unroll
  for (int k = 0; k < D; ++k)
#pragma unroll
    for (int j = 0; j < D; ++j) e[k][j] = x[k + 1][j] - x[0][j];
  if constexpr (D == 3) {
    const G(&a)[3] = e[0];
    const G(&bb)[3] = e[1];
    const G(&cc)[3] = e[2];
    n[0][0] = bb[1] * cc[2] - bb[2] * cc[1];
    n[0][1] = bb[2] * cc[0] - bb[0] * cc[2];
    n[0][2] = bb[0] * cc[1] - bb[1] * cc[0];
    n[1][0] = cc[1] * a[2] - cc[2] * a[1];
    n[1][1] = cc[2] * a[0] - cc[0] * a[2];
    n[1][2] = cc[0] * a[1] - cc[1] * a[0];
    n[2][0] = a[1] * bb[2] - a[2] * bb[1];
    n[2][1] = a[2] * bb[0] - a[0] * bb[2];
    n[2][2] = a[0] * bb[1] - a[1] * bb[0];
    return a[0] * n[0][0] + a[1] * n[0][1] + a[2] * n[0][2];
  } else {
    const G a0 = e[0][0], a1 = e[0][1];
    const G b0 = e[1][0], b1 = e[1][1];
    n[0][0] = b1;  n[0][1] = -b0;
    n[1][0] = -a1; n[1][1] = a0;
    return a0 * b1 - a1 * b0;
  }
}

template <typename G>
__device__ __forceinline__ G abs_g(G v) {
  return v < G(0) ? -v : v;
}

// 1 / a: the hardware reciprocal estimate refined by Newton-Raphson steps
// r <- r + r (1 - a r) (two in double, one in float), within about an ulp of
// the correctly rounded quotient, in place of the IEEE division sequence
// (about half the instructions).
template <typename G>
__device__ __forceinline__ G recip(G a) {
  if constexpr (sizeof(G) == 8) {
    double r = __builtin_amdgcn_rcp(a);
    r = __builtin_fma(r, __builtin_fma(-a, r, 1.0), r);
    return __builtin_fma(r, __builtin_fma(-a, r, 1.0), r);
  } else {
    const float r = __builtin_amdgcn_rcpf(a);
    return __builtin_fmaf(r, __builtin_fmaf(-a, r, 1.0f), r);
  }
}

// Row 0 of the rotated element's local matrix: K[0][p] = scale g_0 . g_p
template <int D, int FORM, typename G>
__device__ __forceinline__ void rot_row(const RotData<D, G>& d, const RowConsts& c, double (&Kr)[D + 1]) {
  constexpr int B = D + 1;
  G e[D][D], n[D][D];
  const G det = simplex_cofactors<D, G>(d.x, e, n);
  if constexpr (sizeof(G) == 4) {
    if (c.redo && tmh::sliver_f32<D>(e, det)) {  // redo this element in double
      RotData<D, double> w;
#pragma unroll
      for (int b = 0; b < B; ++b)
#pragma unroll
        for (int j = 0; j < D; ++j) w.x[b][j] = (double)d.x[b][j];
      w.k = (double)d.k;
      rot_row<D, FORM, double>(w, c, Kr);
      return;
    }
  }
  G coef = (G)c.W;
  if constexpr (FORM == TM_FORM_DIFFUSION_ELEM) coef *= d.k;
  const G scale = coef * recip(abs_g(det));
  // sums start from their first term (a 0 + x the compiler must keep, for
  // the sign of a zero x, costs an instruction each)
  G g0[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    G s = n[0][j];
#pragma unroll
    for (int k = 1; k < D; ++k) s += n[k][j];
    g0[j] = -s;
  }
  G s0 = g0[0] * g0[0];
#pragma unroll
  for (int j = 1; j < D; ++j) s0 += g0[j] * g0[j];
  Kr[0] = (double)(scale * s0);
#pragma unroll
  for (int p = 1; p < B; ++p) {
    G s = g0[0] * n[p - 1][0];
#pragma unroll
    for (int j = 1; j < D; ++j) s += g0[j] * n[p - 1][j];
    Kr[p] = (double)(scale * s);
  }
}

// Row i of the local matrix of the element held in d.
template <int D, int FORM, typename G>
__device__ __forceinline__ void row_from(const ElemData<D, G>& d, const RowConsts& c, int i, double (&Kr)[D + 1]) {
  constexpr int B = D + 1;
  G e[D][D], n[D][D];
  const G det = simplex_cofactors<D, G>(d.x, e, n);
  if constexpr (sizeof(G) == 4) {
    if (c.redo && tmh::sliver_f32<D>(e, det)) {  // redo this element in double
      ElemData<D, double> w;
#pragma unroll
      for (int b = 0; b < B; ++b) {
#pragma unroll
        for (int j = 0; j < D; ++j) w.x[b][j] = (double)d.x[b][j];
        w.k[b] = (double)d.k[b];
      }
      row_from<D, FORM, double>(w, c, i, Kr);
      return;
    }
  }
  const G adet = abs_g(det);
  if constexpr (FORM == TM_FORM_MASS) {
#pragma unroll
    for (int j = 0; j < B; ++j) Kr[j] = (double)(adet * (G)c.mref[i * B + j]);
    return;
  } else {
    G g[B][D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      G s = G(0);
#pragma unroll
      for (int k = 0; k < D; ++k) {
        g[k + 1][j] = n[k][j];
        s += n[k][j];
      }
      g[0][j] = -s;
    }
    G coef = (G)c.W;
    if constexpr (FORM == TM_FORM_DIFFUSION_ELEM) {
      coef *= d.k[0];
    } else if constexpr (FORM == TM_FORM_DIFFUSION_NODE) {
      G kq = G(0);
#pragma unroll
      for (int b = 0; b < B; ++b) kq += d.k[b] * (G)c.wphi[b];
      coef = kq;
    }
    const G scale = coef / adet;
    G gi[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      G s = g[0][j];
#pragma unroll
      for (int b = 1; b < B; ++b) s = (i == b) ? g[b][j] : s;
      gi[j] = s;
    }
#pragma unroll
    for (int j = 0; j < B; ++j) {
      G s = G(0);
#pragma unroll
      for (int k = 0; k < D; ++k) s += gi[k] * g[j][k];
      Kr[j] = (double)(scale * s);
    }
  }
}

// Row i of the local matrix of element e (all three load steps in sequence),
// geometry in double: only runs too large for LDS take this walk, and a
// single geometry type keeps its registers out of the staged paths' budget.
template <typename T, int D, int FORM>
__device__ __forceinline__ void local_row(const int32_t* __restrict__ conn, const T* __restrict__ pts,
                                          const T* __restrict__ coeff, const RowConsts& c, int64_t e, int i,
                                          double (&Kr)[D + 1]) {
  using G = double;
  ElemData<D, G> d;
  load_conn<D, G>(conn, e, d);
  load_geometry<T, D, FORM, G>(pts, coeff, e, d);
  row_from<D, FORM, G>(d, c, i, Kr);
}

// Block bounds: {first row, first incidence, first CSR entry} of blocks blk and
// blk + 1, from the precomputed table (two independent loads) or, without it,
// from blk_ptr -> inc_ptr / row_ptr (two dependent loads).
struct BlockBounds {
  int32_t r0, r1, t0, t1, k0, k1;
};

__device__ __forceinline__ BlockBounds block_bounds(const int32_t* __restrict__ info,
                                                    const int32_t* __restrict__ blk_ptr,
                                                    const int32_t* __restrict__ inc_ptr,
                                                    const int32_t* __restrict__ row_ptr, int64_t blk) {
  BlockBounds b;
  if (info) {
    const int4 a = reinterpret_cast<const int4*>(info)[blk];
    const int4 z = reinterpret_cast<const int4*>(info)[blk + 1];
    b.r0 = a.x; b.t0 = a.y; b.k0 = a.z;
    b.r1 = z.x; b.t1 = z.y; b.k1 = z.z;
  } else {
    b.r0 = blk_ptr[blk];
    b.r1 = blk_ptr[blk + 1];
    b.t0 = inc_ptr[b.r0];
    b.t1 = inc_ptr[b.r1];
    b.k0 = row_ptr[b.r0];
    b.k1 = row_ptr[b.r1];
  }
  return b;
}

// Phase 2 of the row kernels: one thread per CSR entry of the staged rows,
// summing the staged local-row values of the row's incidences in ascending
// element order.  m0 / m1 are the entry masks of the thread's first two
// entries, loaded ahead of phase 1.
// ROT: incidence t's values are staged rotated (s_val[t*B + p] belongs to
// vertex (i + p) mod B, s_pos bytes likewise), so value 0 of every incidence
// of row r is its diagonal contribution: the diagonal entry is a contiguous
// sum over the row's incidences, and only off-diagonal entries walk masks.
template <typename T, int B, bool ROT>
__device__ __forceinline__ void sum_entries(const T* s_val, const uint32_t* s_pos, const int32_t* s_rp,
                                            const int32_t* s_ip, const uint16_t* s_erow,
                                            const uint32_t* __restrict__ entry_mask, int32_t k0, int32_t n_ent,
                                            uint32_t m0, uint32_t m1, T* __restrict__ values) {
  int it = 0;
  for (int32_t kl = threadIdx.x; kl < n_ent; kl += kBlock, ++it) {
    const int32_t lr = s_erow[kl];
    const uint32_t p = (uint32_t)(kl - s_rp[lr]);
    const int32_t ta = s_ip[lr], tz = s_ip[lr + 1];
    T acc = T(0);
    if (ROT && (s_pos[ta] & 0xffu) == p) {
      // diagonal: value 0 of every incidence of the row, in order
      for (int32_t t = ta; t < tz; t += 8) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = t + u < tz ? s_val[(t + u) * B] : T(0);
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (t + u < tz) acc += v[u];
      }
    } else if (entry_mask && tz - ta <= 32) {
      // only the incidences that contribute, lowest bit (= lowest element) first
      uint32_t m = it == 0 ? m0 : it == 1 ? m1 : entry_mask[k0 + kl];
      // four contributions per round: their LDS reads are issued together,
      // the adds stay in ascending element order
      while (m) {
        int32_t q[4];
        bool h[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          h[u] = m != 0u;
          q[u] = ta + (h[u] ? __builtin_ctz(m) : 0);
          m &= m - 1u;
        }
        uint32_t cp[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) cp[u] = s_pos[q[u]];
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          int jj = 0;
#pragma unroll
          for (int j = 1; j < B; ++j) jj = (((cp[u] >> (8 * j)) & 0xffu) == p) ? j : jj;
          v[u] = s_val[q[u] * B + jj];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (h[u]) acc += v[u];
      }
    } else {
      for (int32_t t = ta; t < tz; ++t) {
        const uint32_t cp = s_pos[t];
#pragma unroll
        for (int j = 0; j < B; ++j)
          if (((cp >> (8 * j)) & 0xffu) == p) acc += s_val[t * B + j];
      }
    }
    values[k0 + kl] = acc;
  }
}

// colpos word of incidence (e, i) rotated like the staged values: byte p of
// the result is byte (i + p) mod B of cp.
template <int B>
__device__ __forceinline__ uint32_t rotate_pos(uint32_t cp, int i) {
  uint32_t r = 0;
#pragma unroll
  for (int p = 0; p < B; ++p) {
    const int src = i + p < B ? i + p : i + p - B;
    r |= ((cp >> (8 * src)) & 0xffu) << (8 * p);
  }
  return r;
}

// Destination word of incidence (e, i) rotated like the staged values:
// field p (16 bits) of the result is field (i + p) mod B of w.
template <int B>
__device__ __forceinline__ uint64_t rotate_dest(uint64_t w, int i) {
  if constexpr (B == 4) {
    const int sh = 16 * i;
    return i == 0 ? w : (w >> sh) | (w << (64 - sh));
  } else {
    uint64_t r = 0;
#pragma unroll
    for (int p = 0; p < B; ++p) {
      const int src = i + p < B ? i + p : i + p - B;
      r |= ((w >> (16 * src)) & 0xffffull) << (16 * p);
    }
    return r;
  }
}

// Entry -> local row map of the staged rows (s_rp must be visible).
__device__ __forceinline__ void build_entry_rows(const int32_t* s_rp, int32_t n_rows, uint16_t* s_erow) {
  for (int32_t lr = threadIdx.x; lr < n_rows; lr += kBlock)
    for (int32_t k = s_rp[lr]; k < s_rp[lr + 1]; ++k) s_erow[k] = (uint16_t)lr;
}

// Row / incidence starts of the staged rows (relative to the block) into LDS.
__device__ __forceinline__ void stage_row_maps(const int32_t* __restrict__ row_ptr,
                                               const int32_t* __restrict__ inc_ptr, const BlockBounds& bb,
                                               int32_t* s_rp, int32_t* s_ip) {
  for (int32_t lr = threadIdx.x; lr <= bb.r1 - bb.r0; lr += kBlock) {
    s_rp[lr] = row_ptr[bb.r0 + lr] - bb.k0;
    s_ip[lr] = inc_ptr[bb.r0 + lr] - bb.t0;
  }
}

// Entries [k0, k1) of rows [r0, r1) straight from global memory, each entry
// walking its row's incidences and recomputing the contributing local rows
// (runs too large to stage in LDS).
template <typename T, int D, int FORM>
__device__ void walk_entries(const int32_t* __restrict__ conn, const T* __restrict__ pts,
                             const T* __restrict__ coeff, const RowConsts& c, const int32_t* __restrict__ row_ptr,
                             const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc,
                             const uint32_t* __restrict__ colpos, int32_t r0, int32_t r1, int32_t k0, int32_t k1,
                             T* __restrict__ values) {
  constexpr int B = D + 1;
  for (int32_t k = k0 + threadIdx.x; k < k1; k += kBlock) {
    int32_t lo = r0, hi = r1 - 1;
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (row_ptr[mid] <= k) lo = mid; else hi = mid - 1;
    }
    const int32_t r = lo;
    const uint32_t p = (uint32_t)(k - row_ptr[r]);
    T acc = T(0);
    for (int32_t t = inc_ptr[r]; t < inc_ptr[r + 1]; ++t) {
      const uint32_t cp = colpos[t];
#pragma unroll
      for (int j = 0; j < B; ++j) {
        if (((cp >> (8 * j)) & 0xffu) == p) {
          const int32_t ei = inc[t];
          double Kr[B];
          local_row<T, D, FORM>(conn, pts, coeff, c, ei / B, ei % B, Kr);
          acc += (T)Kr[j];
        }
      }
    }
    values[k] = acc;
  }
}

// The fused kernel with the contribution order of tm_rows_dest_build: phase 1
// writes each local-row value at its position in the workgroup's
// contribution list (entry by entry, ascending element within an entry),
// phase 2 sums every entry's contiguous range.  Its own kernel so that it
// allocates only the LDS it uses (the staged values and the rows' node
// coordinates), which lets more workgroups share a CU.
// Built for seven workgroups per CU (72 VGPRs) in float64 and eight (64
// VGPRs) in float32, except for the nodal coefficient form, which would
// spill: the float32 instances went from 52.7 to 48.0 us at seven and from
// 42.4 to 40.2 us at eight; float64 went 59.9 -> 71.5 us while its hot path
// spilled at seven, fits without spills since the entry ranges are packed,
// and spills again at eight (55 -> 65 us).
template <typename T, int D, int FORM>
__global__ void __launch_bounds__(kBlock, FORM == TM_FORM_DIFFUSION_NODE ? 1 : sizeof(T) == 4 ? 8 : 7)
    p1_rows_dest_kernel(const int32_t* __restrict__ conn, const T* __restrict__ pts, const T* __restrict__ coeff,
                        RowConsts c, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ inc_ptr,
                        const int32_t* __restrict__ inc, const uint32_t* __restrict__ colpos,
                        const int32_t* __restrict__ blk_ptr, const int32_t* __restrict__ blk_info, int64_t n_blocks,
                        const uint64_t* __restrict__ dest, const uint16_t* __restrict__ ent_start,
                        T* __restrict__ values) {
  constexpr int B = D + 1;
  static_assert(kCapInc == 2 * kBlock, "phase 1 stages two incidences per thread");
  using G = Geom<T>;
  __shared__ T s_val[kCapInc * B];
  __shared__ G s_xr[kMaxRows * D];  // coordinates of the staged rows' nodes
  const int64_t blk = xcd_logical(blockIdx.x, n_blocks);
  const BlockBounds bb = block_bounds(blk_info, blk_ptr, inc_ptr, row_ptr, blk);
  const int32_t r0 = bb.r0, r1 = bb.r1, t0 = bb.t0, k0 = bb.k0, k1 = bb.k1;
  const int32_t n_inc = bb.t1 - t0, n_rows = r1 - r0, n_ent = k1 - k0;
  constexpr bool kRot = FORM == TM_FORM_LAPLACE || FORM == TM_FORM_DIFFUSION_ELEM;
  if (n_inc <= kCapInc && n_rows <= kMaxRows && n_ent <= kMaxEnt) {
    // contribution-ordered staging: s_val is the workgroup's contribution
    // list (entry by entry, ascending element within an entry); phase 1
    // writes each local-row value at its list position, phase 2 sums every
    // entry's contiguous range
    T* s_c = s_val;
    if (kRot && (int32_t)threadIdx.x < n_rows) {
      G xr[D];
      load_node<T, D, G>(pts, r0 + (int32_t)threadIdx.x, xr);
#pragma unroll
      for (int j = 0; j < D; ++j) s_xr[threadIdx.x * D + j] = xr[j];
    }
    if (n_inc == 0 || n_ent == 0) return;  // rows of unused nodes only (uniform per workgroup)
    const int32_t ta = threadIdx.x, tb = threadIdx.x + kBlock;
    const bool ha = ta < n_inc, hb = tb < n_inc;
    // every load below is unconditional (indices clamped into the run, a
    // run has at least one incidence and one entry), so that they issue
    // together instead of one exec-masked branch and wait each
    const int32_t ea = inc[t0 + min(ta, n_inc - 1)];
    const int32_t eb = inc[t0 + min(tb, n_inc - 1)];
    const uint64_t wa = dest[t0 + min(ta, n_inc - 1)];
    const uint64_t wb = dest[t0 + min(tb, n_inc - 1)];
    // the thread's entry ranges, loaded ahead of phase 1
    constexpr int kEpt = kMaxEnt / kBlock;
    // (the first two per thread, packed start | end << 16: runs average about
    // 320 entries; the third and fourth are loaded in phase 2 when present)
    constexpr int kEarly = 2;
    uint32_t rng[kEarly];
#pragma unroll
    for (int u = 0; u < kEarly; ++u) {
      const int32_t kl = (int32_t)threadIdx.x + u * kBlock;
      const uint32_t s0 = ent_start[k0 + min(kl, n_ent - 1)];
      const uint32_t s1 = ent_start[k0 + min(kl + 1, n_ent - 1)];
      rng[u] = s0 | ((kl + 1 < n_ent ? s1 : (uint32_t)(n_inc * B)) << 16);
    }
    const int ia = ea % B, ib = eb % B;
    // both incidences' loads are issued unconditionally (element 0 stands in
    // for a missing incidence), so that neither waits for the other's
    const int64_t ga = ea / B, gb = eb / B;
    if constexpr (kRot) {
      RotData<D, G> da, db;
      const int4 wca = load_conn_word<D>(conn, ga);
      const int4 wcb = load_conn_word<D>(conn, gb);
      if constexpr (FORM == TM_FORM_DIFFUSION_ELEM) {
        da.k = (G)coeff[ga];
        db.k = (G)coeff[gb];
      }
      rotate_conn<D, G>(wca, ia, da);
      rotate_conn<D, G>(wcb, ib, db);
      gather_rot<T, D, G>(pts, da);
      gather_rot<T, D, G>(pts, db);
      __syncthreads();  // s_xr
      if (ha) {
#pragma unroll
        for (int j = 0; j < D; ++j) da.x[0][j] = s_xr[(da.c[0] - r0) * D + j];
        double Kr[B];
        rot_row<D, FORM, G>(da, c, Kr);
        const uint64_t w = rotate_dest<B>(wa, ia);
#pragma unroll
        for (int p = 0; p < B; ++p) s_c[(w >> (16 * p)) & 0xffffu] = (T)Kr[p];
      }
      if (hb) {
#pragma unroll
        for (int j = 0; j < D; ++j) db.x[0][j] = s_xr[(db.c[0] - r0) * D + j];
        double Kr[B];
        rot_row<D, FORM, G>(db, c, Kr);
        const uint64_t w = rotate_dest<B>(wb, ib);
#pragma unroll
        for (int p = 0; p < B; ++p) s_c[(w >> (16 * p)) & 0xffffu] = (T)Kr[p];
      }
    } else {
      ElemData<D, G> da, db;
      load_conn<D, G>(conn, ga, da);
      load_conn<D, G>(conn, gb, db);
      load_geometry<T, D, FORM, G>(pts, coeff, ga, da);
      load_geometry<T, D, FORM, G>(pts, coeff, gb, db);
      if (ha) {
        double Kr[B];
        row_from<D, FORM, G>(da, c, ia, Kr);
#pragma unroll
        for (int j = 0; j < B; ++j) s_c[(wa >> (16 * j)) & 0xffffu] = (T)Kr[j];
      }
      if (hb) {
        double Kr[B];
        row_from<D, FORM, G>(db, c, ib, Kr);
#pragma unroll
        for (int j = 0; j < B; ++j) s_c[(wb >> (16 * j)) & 0xffffu] = (T)Kr[j];
      }
    }
    __syncthreads();
    // phase 2: entry kl sums its contributions in list order (ascending
    // element), four LDS reads in flight, the adds in order
#pragma unroll
    for (int u = 0; u < kEpt; ++u) {
      const int32_t kl = (int32_t)threadIdx.x + u * kBlock;
      if (kl < n_ent) {
        int32_t t, te;
        if (u < kEarly) {
          t = (int32_t)(rng[u < kEarly ? u : 0] & 0xffffu);
          te = (int32_t)(rng[u < kEarly ? u : 0] >> 16);
        } else {
          t = ent_start[k0 + kl];
          te = kl + 1 < n_ent ? (int32_t)ent_start[k0 + kl + 1] : n_inc * B;
        }
        T acc = T(0);
        for (; t + 3 < te; t += 4) {
          const T v0 = s_c[t], v1 = s_c[t + 1], v2 = s_c[t + 2], v3 = s_c[t + 3];
          acc += v0;
          acc += v1;
          acc += v2;
          acc += v3;
        }
        for (; t < te; ++t) acc += s_c[t];
        values[k0 + kl] = acc;
      }
    }
    return;
  }
  // oversized run: every entry walks global memory, recomputing local rows
  walk_entries<T, D, FORM>(conn, pts, coeff, c, row_ptr, inc_ptr, inc, colpos, r0, r1, k0, k1, values);
}

template <typename T, int D, int FORM>
__global__ void __launch_bounds__(kBlock)
    p1_rows_kernel(const int32_t* __restrict__ conn, const T* __restrict__ pts, const T* __restrict__ coeff,
                   RowConsts c, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ inc_ptr,
                   const int32_t* __restrict__ inc, const uint32_t* __restrict__ colpos,
                   const uint32_t* __restrict__ entry_mask, const int32_t* __restrict__ blk_ptr,
                   const int32_t* __restrict__ blk_info, int64_t n_blocks, T* __restrict__ values) {
  constexpr int B = D + 1;
  static_assert(kCapInc == 2 * kBlock, "phase 1 stages two incidences per thread");
  __shared__ T s_val[kCapInc * B];
  __shared__ uint32_t s_pos[kCapInc];
  __shared__ int32_t s_rp[kMaxRows + 1];    // row starts, relative to k0
  __shared__ int32_t s_ip[kMaxRows + 1];    // incidence starts, relative to t0
  __shared__ uint16_t s_erow[kMaxEnt];      // entry -> local row
  using G = Geom<T>;
  __shared__ G s_xr[kMaxRows * D];          // coordinates of the staged rows' nodes
  const int64_t blk = xcd_logical(blockIdx.x, n_blocks);
  const BlockBounds bb = block_bounds(blk_info, blk_ptr, inc_ptr, row_ptr, blk);
  const int32_t r0 = bb.r0, r1 = bb.r1, t0 = bb.t0, k0 = bb.k0, k1 = bb.k1;
  const int32_t n_inc = bb.t1 - t0, n_rows = r1 - r0, n_ent = k1 - k0;
  constexpr bool kRot = FORM == TM_FORM_LAPLACE || FORM == TM_FORM_DIFFUSION_ELEM;
  if (n_inc <= kCapInc && n_rows <= kMaxRows && n_ent <= kMaxEnt) {
    // loads that depend only on the block bounds go out first
    stage_row_maps(row_ptr, inc_ptr, bb, s_rp, s_ip);
    if (kRot && (int32_t)threadIdx.x < n_rows) {
      G xr[D];
      load_node<T, D, G>(pts, r0 + (int32_t)threadIdx.x, xr);
#pragma unroll
      for (int j = 0; j < D; ++j) s_xr[threadIdx.x * D + j] = xr[j];
    }
    const int32_t ta = threadIdx.x, tb = threadIdx.x + kBlock;
    const bool ha = ta < n_inc, hb = tb < n_inc;
    uint32_t m0 = 0, m1 = 0;
    if (entry_mask) {
      if (ta < n_ent) m0 = entry_mask[k0 + ta];
      if (tb < n_ent) m1 = entry_mask[k0 + tb];
    }
    // phase 1: the (at most two) incidences of this thread, loads interleaved
    const int32_t ea = ha ? inc[t0 + ta] : 0;
    const int32_t eb = hb ? inc[t0 + tb] : 0;
    const uint32_t pa = ha ? colpos[t0 + ta] : 0u;
    const uint32_t pb = hb ? colpos[t0 + tb] : 0u;
    const int ia = ea % B, ib = eb % B;
    if constexpr (kRot) {
      RotData<D, G> da, db;
      if (ha) load_rot<T, D, FORM, G>(conn, coeff, ea / B, ia, da);
      if (hb) load_rot<T, D, FORM, G>(conn, coeff, eb / B, ib, db);
      if (ha) gather_rot<T, D, G>(pts, da);
      if (hb) gather_rot<T, D, G>(pts, db);
      __syncthreads();  // s_xr, s_rp, s_ip
      build_entry_rows(s_rp, n_rows, s_erow);
      if (ha) {
#pragma unroll
        for (int j = 0; j < D; ++j) da.x[0][j] = s_xr[(da.c[0] - r0) * D + j];
        double Kr[B];
        rot_row<D, FORM, G>(da, c, Kr);
#pragma unroll
        for (int p = 0; p < B; ++p) s_val[ta * B + p] = (T)Kr[p];
        s_pos[ta] = rotate_pos<B>(pa, ia);
      }
      if (hb) {
#pragma unroll
        for (int j = 0; j < D; ++j) db.x[0][j] = s_xr[(db.c[0] - r0) * D + j];
        double Kr[B];
        rot_row<D, FORM, G>(db, c, Kr);
#pragma unroll
        for (int p = 0; p < B; ++p) s_val[tb * B + p] = (T)Kr[p];
        s_pos[tb] = rotate_pos<B>(pb, ib);
      }
    } else {
      ElemData<D, G> da, db;
      if (ha) load_conn<D, G>(conn, ea / B, da);
      if (hb) load_conn<D, G>(conn, eb / B, db);
      if (ha) load_geometry<T, D, FORM, G>(pts, coeff, ea / B, da);
      if (hb) load_geometry<T, D, FORM, G>(pts, coeff, eb / B, db);
      __syncthreads();  // s_rp, s_ip
      build_entry_rows(s_rp, n_rows, s_erow);
      if (ha) {
        double Kr[B];
        row_from<D, FORM, G>(da, c, ia, Kr);
#pragma unroll
        for (int j = 0; j < B; ++j) s_val[ta * B + j] = (T)Kr[j];
        s_pos[ta] = pa;
      }
      if (hb) {
        double Kr[B];
        row_from<D, FORM, G>(db, c, ib, Kr);
#pragma unroll
        for (int j = 0; j < B; ++j) s_val[tb * B + j] = (T)Kr[j];
        s_pos[tb] = pb;
      }
    }
    __syncthreads();
    // phase 2: one thread per entry, incidences in ascending element order
    sum_entries<T, B, kRot>(s_val, s_pos, s_rp, s_ip, s_erow, entry_mask, k0, n_ent, m0, m1, values);
    return;
  }
  // oversized run: every entry walks global memory, recomputing local rows
  walk_entries<T, D, FORM>(conn, pts, coeff, c, row_ptr, inc_ptr, inc, colpos, r0, r1, k0, k1, values);
}

// Row-centric stage S for P1 slot buffers: the same staging and per-entry walk
// as p1_rows_kernel, but phase 1 reads the element rows K_e[i][0..B) from the
// slot buffer written by stage G (one contiguous B-value row per incidence)
// instead of recomputing them.
template <typename T, int D, bool INC_ORDER>
__global__ void __launch_bounds__(kBlock)
    p1_rows_reduce_kernel(const T* __restrict__ slot, const int32_t* __restrict__ row_ptr,
                          const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc,
                          const uint32_t* __restrict__ colpos, const uint32_t* __restrict__ entry_mask,
                          const int32_t* __restrict__ blk_ptr, const int32_t* __restrict__ blk_info,
                          int64_t n_blocks, T* __restrict__ values) {
  constexpr int B = D + 1;
  __shared__ T s_val[kCapInc * B];
  __shared__ uint32_t s_pos[kCapInc];
  __shared__ int32_t s_rp[kMaxRows + 1];
  __shared__ int32_t s_ip[kMaxRows + 1];
  __shared__ uint16_t s_erow[kMaxEnt];
  const int64_t blk = xcd_logical(blockIdx.x, n_blocks);
  const BlockBounds bb = block_bounds(blk_info, blk_ptr, inc_ptr, row_ptr, blk);
  const int32_t r0 = bb.r0, r1 = bb.r1, t0 = bb.t0, k0 = bb.k0, k1 = bb.k1;
  const int32_t n_inc = bb.t1 - t0, n_rows = r1 - r0, n_ent = k1 - k0;
  if (n_inc <= kCapInc && n_rows <= kMaxRows && n_ent <= kMaxEnt) {
    stage_row_maps(row_ptr, inc_ptr, bb, s_rp, s_ip);
    uint32_t m0 = 0, m1 = 0;
    if (entry_mask) {
      if ((int32_t)threadIdx.x < n_ent) m0 = entry_mask[k0 + threadIdx.x];
      if ((int32_t)threadIdx.x + kBlock < n_ent) m1 = entry_mask[k0 + threadIdx.x + kBlock];
    }
    for (int32_t t = threadIdx.x; t < n_inc; t += kBlock) {
      // element-ordered slots: row (e, i) starts at (e*B + i)*B; incidence-ordered: at (t0 + t)*B
      const T* src = INC_ORDER ? slot + (int64_t)(t0 + t) * B : slot + (int64_t)inc[t0 + t] * B;
      if constexpr (B == 4 && sizeof(T) == 8) {
        const double2 a = reinterpret_cast<const double2*>(src)[0];
        const double2 b = reinterpret_cast<const double2*>(src)[1];
        s_val[t * B + 0] = a.x; s_val[t * B + 1] = a.y;
        s_val[t * B + 2] = b.x; s_val[t * B + 3] = b.y;
      } else if constexpr (B == 4 && sizeof(T) == 4) {
        const float4 a = reinterpret_cast<const float4*>(src)[0];
        s_val[t * B + 0] = a.x; s_val[t * B + 1] = a.y;
        s_val[t * B + 2] = a.z; s_val[t * B + 3] = a.w;
      } else {
#pragma unroll
        for (int j = 0; j < B; ++j) s_val[t * B + j] = src[j];
      }
      s_pos[t] = colpos[t0 + t];
    }
    __syncthreads();
    for (int32_t lr = threadIdx.x; lr < n_rows; lr += kBlock)
      for (int32_t k = s_rp[lr]; k < s_rp[lr + 1]; ++k) s_erow[k] = (uint16_t)lr;
    __syncthreads();
    sum_entries<T, B, false>(s_val, s_pos, s_rp, s_ip, s_erow, entry_mask, k0, n_ent, m0, m1, values);
    return;
  }
  for (int32_t k = k0 + threadIdx.x; k < k1; k += kBlock) {
    int32_t lo = r0, hi = r1 - 1;
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (row_ptr[mid] <= k) lo = mid; else hi = mid - 1;
    }
    const int32_t r = lo;
    const uint32_t p = (uint32_t)(k - row_ptr[r]);
    T acc = T(0);
    for (int32_t t = inc_ptr[r]; t < inc_ptr[r + 1]; ++t) {
      const uint32_t cp = colpos[t];
#pragma unroll
      for (int j = 0; j < B; ++j)
        if (((cp >> (8 * j)) & 0xffu) == p) acc += slot[(INC_ORDER ? (int64_t)t : (int64_t)inc[t]) * B + j];
    }
    values[k] = acc;
  }
}

template <typename T, int D>
int launch(const tm_mesh_view* m, const tm_form_desc* f, const tm_rows_view* rv, void* values, hipStream_t s) {
  constexpr int B = D + 1;
  RowConsts c{};
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
  c.redo = (f->flags & TM_FLAG_F32_NO_REDO) ? 0 : 1;
  const T* pts = static_cast<const T*>(m->points);
  const T* coeff = static_cast<const T*>(f->coeff);
  T* out = static_cast<T*>(values);
  const int64_t nb = rv->n_blocks;
  if (nb == 0) return TM_OK;
  const uint64_t* dest = (rv->dest && rv->ent_start) ? reinterpret_cast<const uint64_t*>(rv->dest) : nullptr;
#define TM_ROWS_LAUNCH(FORM)                                                                                \
  do {                                                                                                      \
    if (dest)                                                                                               \
      p1_rows_dest_kernel<T, D, FORM><<<nb, kBlock, 0, s>>>(m->conn, pts, coeff, c, rv->row_ptr, rv->inc_ptr, \
                                                            rv->inc, rv->colpos, rv->blk_ptr, rv->blk_info, nb, \
                                                            dest, rv->ent_start, out);                     \
    else                                                                                                    \
      p1_rows_kernel<T, D, FORM><<<nb, kBlock, 0, s>>>(m->conn, pts, coeff, c, rv->row_ptr, rv->inc_ptr,       \
                                                       rv->inc, rv->colpos, rv->entry_mask, rv->blk_ptr,        \
                                                       rv->blk_info, nb, out);                                  \
  } while (0)
  switch (f->form) {
    case TM_FORM_LAPLACE: TM_ROWS_LAUNCH(TM_FORM_LAPLACE); break;
    case TM_FORM_DIFFUSION_ELEM:
      if (!coeff) return tmh::fail(TM_ERR_ARG, "diffusion form needs a coefficient");
      TM_ROWS_LAUNCH(TM_FORM_DIFFUSION_ELEM);
      break;
    case TM_FORM_DIFFUSION_NODE:
      if (!coeff) return tmh::fail(TM_ERR_ARG, "diffusion form needs a coefficient");
      TM_ROWS_LAUNCH(TM_FORM_DIFFUSION_NODE);
      break;
    case TM_FORM_MASS: TM_ROWS_LAUNCH(TM_FORM_MASS); break;
    default: return tmh::fail(TM_ERR_UNSUPPORTED, "form not available for the fused P1 path");
  }
#undef TM_ROWS_LAUNCH
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("p1 rows kernel launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
}

}  // namespace

extern "C" int tm_assemble_p1_rows(const tm_mesh_view* mesh, const tm_form_desc* form, const tm_rows_view* rows,
                                   void* values, tm_stream_t stream_) {
  TM_TRY_BEGIN
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  if (!mesh || !form || !rows) return tmh::fail(TM_ERR_ARG, "NULL argument");
  if (mesh->cell != TM_TRI3 && mesh->cell != TM_TET4) return tmh::fail(TM_ERR_UNSUPPORTED, "fused path is P1 only");
  if (mesh->dtype != TM_F32 && mesh->dtype != TM_F64) return tmh::fail(TM_ERR_ARG, "unknown dtype");
  if (mesh->n_elem * 4 >= tmh::kMaxIndex || rows->nnz >= tmh::kMaxIndex) return tmh::fail(TM_ERR_SIZE, "too large");
  if (rows->n_blocks > 0 && (!rows->row_ptr || !rows->inc_ptr || !rows->inc || !rows->colpos || !rows->blk_ptr ||
                             !values || !mesh->conn || !mesh->points))
    return tmh::fail(TM_ERR_ARG, "NULL buffer");
  if (mesh->dtype == TM_F64)
    return mesh->cell == TM_TET4 ? launch<double, 3>(mesh, form, rows, values, stream)
                                 : launch<double, 2>(mesh, form, rows, values, stream);
  return mesh->cell == TM_TET4 ? launch<float, 3>(mesh, form, rows, values, stream)
                               : launch<float, 2>(mesh, form, rows, values, stream);
  TM_TRY_END
}

int tm_rows_reduce_launch(int32_t dtype, int32_t cell, const tm_rows_view* rows, const void* slot, bool inc_order,
                          void* values, hipStream_t stream) {
  if (!rows) return tmh::fail(TM_ERR_ARG, "NULL rows view");
  if (cell != TM_TRI3 && cell != TM_TET4) return tmh::fail(TM_ERR_UNSUPPORTED, "row reduction is P1 only");
  const int64_t nb = rows->n_blocks;
  if (nb == 0) return TM_OK;
  if (!slot || !values || !rows->row_ptr || !rows->inc_ptr || !rows->inc || !rows->colpos || !rows->blk_ptr)
    return tmh::fail(TM_ERR_ARG, "NULL buffer");
#define TM_RR(T, D, IO)                                                                                          \
  p1_rows_reduce_kernel<T, D, IO><<<nb, kBlock, 0, stream>>>((const T*)slot, rows->row_ptr, rows->inc_ptr, rows->inc, \
                                                             rows->colpos, rows->entry_mask, rows->blk_ptr,          \
                                                             rows->blk_info, nb, (T*)values)
  if (dtype == TM_F64) {
    if (cell == TM_TET4) { if (inc_order) TM_RR(double, 3, true); else TM_RR(double, 3, false); }
    else { if (inc_order) TM_RR(double, 2, true); else TM_RR(double, 2, false); }
  } else if (dtype == TM_F32) {
    if (cell == TM_TET4) { if (inc_order) TM_RR(float, 3, true); else TM_RR(float, 3, false); }
    else { if (inc_order) TM_RR(float, 2, true); else TM_RR(float, 2, false); }
  } else {
    return tmh::fail(TM_ERR_ARG, "unknown dtype");
  }
#undef TM_RR
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("row reduction launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
}

extern "C" int tm_reduce_p1_rows(int32_t dtype, int32_t cell, const tm_rows_view* rows, const void* slot,
                                 int32_t inc_order, void* values, tm_stream_t stream_) {
  TM_TRY_BEGIN
  return tm_rows_reduce_launch(dtype, cell, rows, slot, inc_order != 0, values,
                               reinterpret_cast<hipStream_t>(stream_));
  TM_TRY_END
}

namespace {

// One workgroup per plan block: every row's incidences, in order, number
// their contributions per entry (row-local ranks), the workgroup scans the
// per-entry counts into entry starts, and each incidence's ranks become list
// positions.  Blocks over the staging caps are left alone.
__global__ void __launch_bounds__(kBlock)
    rows_dest_kernel(int B, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ inc_ptr,
                     const uint32_t* __restrict__ colpos, const int32_t* __restrict__ blk_ptr,
                     const int32_t* __restrict__ blk_info, int64_t n_blocks, uint16_t* __restrict__ dest,
                     uint16_t* __restrict__ ent_start) {
  __shared__ uint16_t s_cnt[kMaxEnt];
  __shared__ uint16_t s_start[kMaxEnt];
  const int64_t blk = blockIdx.x;
  const BlockBounds bb = block_bounds(blk_info, blk_ptr, inc_ptr, row_ptr, blk);
  const int32_t n_inc = bb.t1 - bb.t0, n_rows = bb.r1 - bb.r0, n_ent = bb.k1 - bb.k0;
  if (!(n_inc <= kCapInc && n_rows <= kMaxRows && n_ent <= kMaxEnt)) return;  // uniform per workgroup
  for (int32_t k = threadIdx.x; k < n_ent; k += kBlock) s_cnt[k] = 0;
  __syncthreads();
  // ranks within each entry, one thread per row (rows own disjoint entries)
  for (int32_t lr = threadIdx.x; lr < n_rows; lr += kBlock) {
    const int32_t r = bb.r0 + lr;
    const int32_t kr = row_ptr[r] - bb.k0;
    for (int32_t t = inc_ptr[r]; t < inc_ptr[r + 1]; ++t) {
      const uint32_t cp = colpos[t];
      for (int j = 0; j < 4; ++j) {
        uint16_t rank = 0;
        if (j < B) {
          const int32_t kl = kr + (int32_t)((cp >> (8 * j)) & 0xffu);
          rank = s_cnt[kl];
          s_cnt[kl] = rank + 1;
        }
        dest[(int64_t)t * 4 + j] = rank;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int32_t k = 0; k < n_ent; ++k) {
      s_start[k] = (uint16_t)run;
      run += s_cnt[k];
    }
  }
  __syncthreads();
  for (int32_t k = threadIdx.x; k < n_ent; k += kBlock) ent_start[bb.k0 + k] = s_start[k];
  for (int32_t lr = threadIdx.x; lr < n_rows; lr += kBlock) {
    const int32_t r = bb.r0 + lr;
    const int32_t kr = row_ptr[r] - bb.k0;
    for (int32_t t = inc_ptr[r]; t < inc_ptr[r + 1]; ++t) {
      const uint32_t cp = colpos[t];
      for (int j = 0; j < B; ++j) {
        const int32_t kl = kr + (int32_t)((cp >> (8 * j)) & 0xffu);
        dest[(int64_t)t * 4 + j] = (uint16_t)(dest[(int64_t)t * 4 + j] + s_start[kl]);
      }
    }
  }
}

}  // namespace

extern "C" int tm_rows_dest_build(int32_t n_basis, const tm_rows_view* rows, uint16_t* dest, uint16_t* ent_start,
                                  tm_stream_t stream_) {
  TM_TRY_BEGIN
  if (!rows) return tmh::fail(TM_ERR_ARG, "NULL rows view");
  if (n_basis != 3 && n_basis != 4) return tmh::fail(TM_ERR_UNSUPPORTED, "contribution order is for P1 cells");
  const int64_t nb = rows->n_blocks;
  if (nb == 0) return TM_OK;
  if (!rows->row_ptr || !rows->inc_ptr || !rows->colpos || !dest || !ent_start || (!rows->blk_ptr && !rows->blk_info))
    return tmh::fail(TM_ERR_ARG, "NULL buffer");
  rows_dest_kernel<<<nb, kBlock, 0, reinterpret_cast<hipStream_t>(stream_)>>>(
      n_basis, rows->row_ptr, rows->inc_ptr, rows->colpos, rows->blk_ptr, rows->blk_info, nb, dest, ent_start);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("contribution order launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
  TM_TRY_END
}
