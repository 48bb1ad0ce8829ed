// Block-row reduction of incidence-ordered element rows, and the Neo-Hookean
// residual / tangent assembled through it.
//
// The reference sums block-form local matrices with ReduceProjector
// (projector.py:95-123, index_add_ over slots e*B^2 + a*B + b) into the BSR
// layout of from_block_coo (matrix.py:331-381).  Here the element kernel
// (hyper_wave_kernel in tm_p2.hip) writes the block row (e, a) -- the B
// blocks of H x H values of node conn[e, a] -- at its position in the
// node-sorted incidence list, so the contributions of every block row of the
// matrix form one contiguous run of slot rows.  block_rows_reduce_kernel
// gives a workgroup a run of whole block rows:
//   * in chunks of incidences, it copies their slot rows into LDS with
//     coalesced loads (the next chunk's rows are fetched into registers while
//     the current one is summed);
//   * one thread per (block entry, component) adds the staged blocks of its
//     contributions in ascending element order, following a per-entry list of
//     workgroup-local (incidence, column) pairs built once per pattern
//     (block_pairs_kernel) -- the order of the reference's index_add_;
//   * the residual rows (H values per incidence) are summed per node over its
//     incidences in the same order.
// The element tangent itself is computed once per element (closed form, see
// tm_p2.hip); compared with the segment reduction of tm_reduce.hip this reads
// each slot row once, contiguously, instead of gathering H*H values per slot.
#include <hip/hip_runtime.h>

#include "tm_internal.h"

int tm_hyper_element_rows(const tm_mesh_view* mesh, const tm_form_desc* form, const void* disp, void* tslot,
                          void* rslot, const int32_t* inc_inv, hipStream_t stream);

namespace {

constexpr int kBlk = 256;
constexpr int kSlotsPer = 4;     // (entry, component) accumulators per thread
constexpr int kMaxEC = kBlk * kSlotsPer;
constexpr int kStageVals = 2304;  // slot values staged per chunk
constexpr int kPre = kStageVals / kBlk;
constexpr int kMaxPairs = 1536;  // contributions of a workgroup staged in LDS

__device__ __forceinline__ int64_t xcd_logical(int64_t b, int64_t nb) {
  const int64_t q = nb / 8, r = nb % 8;
  const int64_t x = b % 8, y = b / 8;
  return x < r ? x * (q + 1) + y : r * (q + 1) + (x - r) * q + y;
}

// B basis functions, HH values per block, RD residual values per incidence
template <typename T, int B, int HH, int RD>
__global__ void __launch_bounds__(kBlk)
    block_rows_reduce_kernel(const T* __restrict__ tslot, const T* __restrict__ rslot,
                             const int32_t* __restrict__ seg_ptr, const uint16_t* __restrict__ pairs,
                             const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ blk_info,
                             int64_t n_blocks, T* __restrict__ tangent, T* __restrict__ residual) {
  constexpr int kRow = B * HH;                // values per slot row
  constexpr int kChunk = kStageVals / kRow;   // incidences per chunk
  static_assert(kChunk >= 1, "a slot row must fit one chunk");
  __shared__ T s_val[kChunk * kRow];
  __shared__ T s_res[kChunk * RD];
  __shared__ uint16_t s_pairs[kMaxPairs];
  const int64_t blk = xcd_logical(blockIdx.x, n_blocks);
  const int4 lo = reinterpret_cast<const int4*>(blk_info)[blk];
  const int4 hi = reinterpret_cast<const int4*>(blk_info)[blk + 1];
  const int32_t r0 = lo.x, t0 = lo.y, k0 = lo.z;
  const int32_t n_rows = hi.x - r0, n_inc = hi.y - t0, n_ent = hi.z - k0;
  const int32_t p0 = lo.w, p1 = hi.w;  // the workgroup's run of contributions
  const bool lds_pairs = p1 - p0 <= kMaxPairs;
  if (lds_pairs)
    for (int32_t c = threadIdx.x; c < p1 - p0; c += kBlk) s_pairs[c] = pairs[p0 + c];
  const int32_t n_chunks = (n_inc + kChunk - 1) / kChunk;
  // slot rows of chunk ci: values [(t0 + ci*kChunk) * kRow, ...), contiguous
  auto fetch = [&](int32_t ci, T (&pre)[kPre]) {
    const int32_t c0 = ci * kChunk;
    const int32_t nv = (n_inc - c0 < kChunk ? n_inc - c0 : kChunk) * kRow;
    const T* src = tslot + ((int64_t)t0 + c0) * kRow;
#pragma unroll
    for (int j = 0; j < kPre; ++j) {
      const int idx = threadIdx.x + j * kBlk;
      pre[j] = (tslot && idx < nv) ? src[idx] : T(0);
    }
  };
  const int64_t n_ec = (int64_t)n_ent * HH;
  // one pass (more only for very long rows); always one, so nodes without
  // incidences still get their zero residual
  for (int64_t base = 0; base == 0 || base < n_ec; base += kMaxEC) {
    T acc[kSlotsPer];
    int32_t ptr[kSlotsPer], end[kSlotsPer];
#pragma unroll
    for (int j = 0; j < kSlotsPer; ++j) {
      acc[j] = T(0);
      const int64_t s = base + threadIdx.x + (int64_t)j * kBlk;
      if (s < n_ec) {
        const int32_t kl = (int32_t)(s / HH);
        ptr[j] = seg_ptr[k0 + kl] - p0;
        end[j] = seg_ptr[k0 + kl + 1] - p0;
      } else {
        ptr[j] = 0;
        end[j] = 0;
      }
    }
    const bool has_res = residual && base == 0 && (int32_t)threadIdx.x < n_rows * RD;
    const int32_t rr = has_res ? (int32_t)threadIdx.x / RD : 0;
    int32_t rptr = has_res ? inc_ptr[r0 + rr] - t0 : 0;
    const int32_t rend = has_res ? inc_ptr[r0 + rr + 1] - t0 : 0;
    T racc = T(0);
    T pre[kPre];
    if (n_chunks > 0) fetch(0, pre);
    for (int32_t ci = 0; ci < n_chunks; ++ci) {
      const int32_t c0 = ci * kChunk;
      const int32_t nc = n_inc - c0 < kChunk ? n_inc - c0 : kChunk;
      __syncthreads();  // readers of the previous chunk are done (first time: s_pairs)
#pragma unroll
      for (int j = 0; j < kPre; ++j) {
        const int idx = threadIdx.x + j * kBlk;
        if (idx < kChunk * kRow) s_val[idx] = pre[j];
      }
      if (rslot)
        for (int idx = threadIdx.x; idx < nc * RD; idx += kBlk) s_res[idx] = rslot[((int64_t)t0 + c0) * RD + idx];
      __syncthreads();
      if (ci + 1 < n_chunks) fetch(ci + 1, pre);  // in flight while this chunk is summed
#pragma unroll
      for (int j = 0; j < kSlotsPer; ++j) {
        const int comp = (int)((base + threadIdx.x + (int64_t)j * kBlk) % HH);
        while (ptr[j] < end[j]) {
          const int32_t pr = lds_pairs ? s_pairs[ptr[j]] : pairs[p0 + ptr[j]];
          const int32_t pt = pr / B;
          if (pt >= c0 + nc) break;
          acc[j] += s_val[(pt - c0) * kRow + (pr % B) * HH + comp];
          ++ptr[j];
        }
      }
      if (has_res) {
        const int comp = (int)threadIdx.x % RD;
        while (rptr < rend && rptr < c0 + nc) {
          racc += s_res[(rptr - c0) * RD + comp];
          ++rptr;
        }
      }
    }
    if (tangent) {
#pragma unroll
      for (int j = 0; j < kSlotsPer; ++j) {
        const int64_t s = base + threadIdx.x + (int64_t)j * kBlk;
        if (s < n_ec) tangent[(int64_t)k0 * HH + s] = acc[j];
      }
    }
    if (has_res) residual[(int64_t)(r0 + rr) * RD + (int)threadIdx.x % RD] = racc;
  }
}

// pairs[c] for contribution c of entry k of row r: slot s = seg_slot[c] =
// e*B^2 + a*B + b, (e, a) at position inc_inv[e*B + a] of the node-sorted
// incidence list; workgroup-local pair = (inc_inv - first incidence of r's
// workgroup) * B + b.
__global__ void __launch_bounds__(256)
    block_pairs_kernel(int32_t B, int64_t n_points, const int32_t* __restrict__ row_ptr,
                       const int32_t* __restrict__ seg_ptr, const int32_t* __restrict__ seg_slot,
                       const int32_t* __restrict__ inc_inv, const int32_t* __restrict__ row_t0,
                       uint16_t* __restrict__ pairs, int32_t* __restrict__ overflow) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= n_points) return;
  const int32_t base = row_t0[r];
  for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
    for (int32_t c = seg_ptr[k]; c < seg_ptr[k + 1]; ++c) {
      const int32_t s = seg_slot[c];
      const int32_t p = (inc_inv[s / B] - base) * B + s % B;
      if (p < 0 || p > 0xffff) overflow[0] = 1;
      pairs[c] = (uint16_t)(p & 0xffff);
    }
}

template <typename T, int B, int HH, int RD>
int launch_rows_reduce(const tm_block_rows_view* v, const void* tslot, void* tangent, const void* rslot,
                       void* residual, hipStream_t s) {
  if (v->n_blocks <= 0) return TM_OK;
  block_rows_reduce_kernel<T, B, HH, RD><<<v->n_blocks, kBlk, 0, s>>>(
      static_cast<const T*>(tslot), static_cast<const T*>(rslot), v->seg_ptr, v->pairs, v->inc_ptr, v->blk_info,
      v->n_blocks, static_cast<T*>(tangent), static_cast<T*>(residual));
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("block row reduction launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
}

template <typename T>
int dispatch_rows_reduce(int32_t cell, const tm_block_rows_view* v, const void* tslot, void* tangent,
                         const void* rslot, void* residual, hipStream_t s) {
  switch (cell) {
    case TM_TRI3: return launch_rows_reduce<T, 3, 4, 2>(v, tslot, tangent, rslot, residual, s);
    case TM_TRI6: return launch_rows_reduce<T, 6, 4, 2>(v, tslot, tangent, rslot, residual, s);
    case TM_TET4: return launch_rows_reduce<T, 4, 9, 3>(v, tslot, tangent, rslot, residual, s);
    case TM_TET10: return launch_rows_reduce<T, 10, 9, 3>(v, tslot, tangent, rslot, residual, s);
    default: return tmh::fail(TM_ERR_UNSUPPORTED, "cell type");
  }
}

}  // namespace

extern "C" int tm_block_rows_pairs(int32_t n_basis, int64_t n_points, const int32_t* row_ptr, const int32_t* seg_ptr,
                                   const int32_t* seg_slot, const int32_t* inc_inv, const int32_t* row_t0,
                                   uint16_t* pairs, int32_t* overflow, tm_stream_t stream_) {
  TM_TRY_BEGIN
  if (!row_ptr || !seg_ptr || !seg_slot || !inc_inv || !row_t0 || !pairs || !overflow)
    return tmh::fail(TM_ERR_ARG, "NULL argument");
  if (n_basis <= 0 || n_basis > 10 || n_points < 0 || n_points >= tmh::kMaxIndex)
    return tmh::fail(TM_ERR_ARG, "bad sizes");
  if (n_points == 0) return TM_OK;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream_);
  block_pairs_kernel<<<(n_points + 255) / 256, 256, 0, s>>>(n_basis, n_points, row_ptr, seg_ptr, seg_slot, inc_inv,
                                                           row_t0, pairs, overflow);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("block row pairs launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
  TM_TRY_END
}

extern "C" int tm_assemble_neohookean_rows(const tm_mesh_view* mesh, const tm_form_desc* form, const void* disp,
                                           const tm_block_rows_view* rows, const int32_t* inc_inv, void* tslot_ws,
                                           void* tangent, void* rslot_ws, void* residual, tm_stream_t stream_) {
  TM_TRY_BEGIN
  hipStream_t s = reinterpret_cast<hipStream_t>(stream_);
  if (!mesh || !form || !rows || !disp || !inc_inv) return tmh::fail(TM_ERR_ARG, "NULL argument");
  if (form->flags & TM_FLAG_UPPER)
    return tmh::fail(TM_ERR_ARG, "the block-row reduction reads whole element rows: TM_FLAG_UPPER is not accepted");
  if (!tangent && !residual) return TM_OK;
  if ((tangent && !tslot_ws) || (residual && !rslot_ws)) return tmh::fail(TM_ERR_ARG, "slot workspace missing");
  if (rows->n_blocks > 0 && (!rows->seg_ptr || !rows->pairs || !rows->inc_ptr || !rows->blk_info))
    return tmh::fail(TM_ERR_ARG, "NULL buffer");
  const int64_t B = tmh::cell_basis(mesh->cell);
  if (B < 0) return tmh::fail(TM_ERR_ARG, "unknown cell");
  if (mesh->n_elem * B * B >= tmh::kMaxIndex) return tmh::fail(TM_ERR_SIZE, "mesh too large");
  if (mesh->dtype != TM_F64 && mesh->dtype != TM_F32) return tmh::fail(TM_ERR_ARG, "unknown dtype");
  void* ts = tangent ? tslot_ws : nullptr;
  void* rs = residual ? rslot_ws : nullptr;
  int st = tm_hyper_element_rows(mesh, form, disp, ts, rs, inc_inv, s);
  if (st) return st;
  return mesh->dtype == TM_F64 ? dispatch_rows_reduce<double>(mesh->cell, rows, ts, tangent, rs, residual, s)
                               : dispatch_rows_reduce<float>(mesh->cell, rows, ts, tangent, rs, residual, s);
  TM_TRY_END
}
