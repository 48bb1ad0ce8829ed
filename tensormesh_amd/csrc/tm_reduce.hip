// Stage S: deterministic segmented reduction of slot values into entries.
//
// Replaces ReduceProjector.__call__ (tensormesh/assemble/projector.py:95-123),
// whose index_add_ uses float atomics on a GPU.  Entry k is the sum of
// slot[seg_slot[t]] for t in [seg_ptr[k], seg_ptr[k+1]), added from zero left to
// right -- the ascending-slot order in which the reference's sequential CPU
// index_add_ accumulates -- so results are bitwise reproducible.
//
// With a reduction plan (tm_reduce_plan) each workgroup owns a run of whole
// entries whose contributions fit in LDS ("CSR-stream" SpMV with unit values):
//   1. all 256 threads gather the run's contributions into LDS, reading
//      seg_slot contiguously and keeping many independent gathers in flight;
//   2. one thread per entry sums its contiguous LDS range in order and writes
//      the result (coalesced across the workgroup).
// Runs that exceed the LDS capacity (a single very long entry) fall back to a
// thread-per-entry walk over global memory.
//
// Workgroups of gathered runs are remapped so that each XCD gets a
// contiguous range of runs: neighbouring runs read the same element slot
// rows, which then stay in one XCD's L2.  Contiguous runs (contribution-
// ordered slots) run in descending order instead (see the kernel).
//
// Symmetric segment lists (tm_symmetric_segments): slot ids name packed
// upper blocks (element kernels of symmetric forms store only the blocks
// i <= j, contiguously per element), and a slot id with bit 31 set names the
// transposed block -- value h = a*D + b of the entry is read from position
// b*D + a of the slot.
//
// Paired outputs (tm_segmented_reduce_pairs): segment k is summed once and
// written to entry[k] and, transposed, to entry_t[k] -- the upper entries
// (r <= c) of a symmetric matrix fill the lower ones, so every slot block is
// read once instead of twice.
#include <hip/hip_runtime.h>

#include "tm_internal.h"

namespace {

constexpr int kBlock = 256;
// LDS chunk per workgroup: 16 KB measured faster than 8, 32 and 64 KB (the
// gathers are latency-bound; more, smaller workgroups keep more in flight)
constexpr int kLdsBytes = 16384;
constexpr int32_t kTransposed = INT32_MIN;  // bit 31 of a slot id

// Value index inside a slot for value h of an entry: transposed for D x D
// blocks (HH = 4, 9) when the slot id carries bit 31; other HH ignore it.
template <int HH>
__device__ __forceinline__ int64_t slot_value(int32_t raw, int h) {
  constexpr int kD = HH == 9 ? 3 : HH == 4 ? 2 : 1;
  const int64_t s = raw & INT32_MAX;
  const int hv = (kD > 1 && raw < 0) ? (h % kD) * kD + h / kD : h;
  return s * HH + hv;
}

// Writes the sum of segment k: out[k] (entry == NULL), or out[entry[k]] and,
// transposed for D x D blocks, out[entry_t[k]] when that is another entry.
template <typename T, int HH>
__device__ __forceinline__ void write_entry(T* __restrict__ out, int64_t k, const T (&acc)[HH],
                                            const int32_t* __restrict__ entry, const int32_t* __restrict__ entry_t) {
  if (!entry) {
#pragma unroll
    for (int h = 0; h < HH; ++h) out[k * HH + h] = acc[h];
    return;
  }
  constexpr int kD = HH == 9 ? 3 : HH == 4 ? 2 : 1;
  const int64_t k0 = entry[k], k1 = entry_t[k];
#pragma unroll
  for (int h = 0; h < HH; ++h) out[k0 * HH + h] = acc[h];
  if (k1 != k0) {
#pragma unroll
    for (int h = 0; h < HH; ++h) out[k1 * HH + (h % kD) * kD + h / kD] = acc[h];
  }
}

// Bijection from the hardware dispatch order (block b runs on XCD b % 8) to a
// logical order in which every XCD owns one contiguous range.
__device__ __forceinline__ int64_t xcd_logical(int64_t b, int64_t nb) {
  const int64_t q = nb / 8, r = nb % 8;
  const int64_t x = b % 8, y = b / 8;
  return x < r ? x * (q + 1) + y : r * (q + 1) + (x - r) * q + y;
}

// seg_slot == NULL: contiguous contributions (contribution t is slot t)
template <typename T, int HH>
__global__ void __launch_bounds__(kBlock) reduce_walk_kernel(const int32_t* __restrict__ seg_ptr,
                                                             const int32_t* __restrict__ seg_slot,
                                                             const T* __restrict__ slot, int64_t nnz,
                                                             const int32_t* __restrict__ entry,
                                                             const int32_t* __restrict__ entry_t,
                                                             T* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= nnz) return;
  T acc[HH];
#pragma unroll
  for (int h = 0; h < HH; ++h) acc[h] = T(0);
  for (int32_t t = seg_ptr[k]; t < seg_ptr[k + 1]; ++t) {
    const int32_t raw = seg_slot ? seg_slot[t] : t;
#pragma unroll
    for (int h = 0; h < HH; ++h) acc[h] += slot[slot_value<HH>(raw, h)];
  }
  write_entry<T, HH>(out, k, acc, entry, entry_t);
}

template <typename T, int HH>
__global__ void __launch_bounds__(kBlock) reduce_stream_kernel(const int32_t* __restrict__ blk_ptr, int64_t n_blocks,
                                                               const int32_t* __restrict__ seg_ptr,
                                                               const int32_t* __restrict__ seg_slot,
                                                               const T* __restrict__ slot,
                                                               const int32_t* __restrict__ entry,
                                                               const int32_t* __restrict__ entry_t,
                                                               T* __restrict__ out) {
  constexpr int kCap = kLdsBytes / (int)(sizeof(T) * HH) < 4096 ? kLdsBytes / (int)(sizeof(T) * HH) : 4096;
  __shared__ T lds[kCap * HH];
  // gathered runs: XCD-contiguous ranges (neighbouring runs gather the same
  // element slots); contiguous runs share nothing and go in descending order:
  // the element kernels write slots roughly in ascending row order, so the
  // slots written last, the likeliest to be still cached, are read first
  // (measured: Neo-Hookean f64 406 -> 380 us, hexahedral elasticity
  // 304 -> 287 us)
  const int64_t b = seg_slot ? xcd_logical(blockIdx.x, n_blocks) : n_blocks - 1 - (int64_t)blockIdx.x;
  const int32_t kb = blk_ptr[b], ke = blk_ptr[b + 1];
  const int32_t cb = seg_ptr[kb], ce = seg_ptr[ke];
  const int32_t n = ce - cb;
  if (n > kCap) {
    // one oversized entry run: walk global memory
    for (int32_t k = kb + threadIdx.x; k < ke; k += kBlock) {
      T acc[HH];
#pragma unroll
      for (int h = 0; h < HH; ++h) acc[h] = T(0);
      for (int32_t t = seg_ptr[k]; t < seg_ptr[k + 1]; ++t) {
        const int32_t raw = seg_slot ? seg_slot[t] : t;
#pragma unroll
        for (int h = 0; h < HH; ++h) acc[h] += slot[slot_value<HH>(raw, h)];
      }
      write_entry<T, HH>(out, k, acc, entry, entry_t);
    }
    return;
  }
  // phase-2 lane mapping (below), decided and its loads issued up front:
  // they depend only on the run's bounds, not on the staged values
  constexpr int kD = HH == 9 ? 3 : HH == 4 ? 2 : 1;
  constexpr int L = kD > 1 ? kD : HH;  // lanes per entry in the lane-group mode
  constexpr int kPerPass = kBlock / L;
  // more than one pass per lane group measured slower for float32 slots
  // (whose 16 KB runs hold twice the entries) than a thread per entry
  constexpr int kMaxPass = sizeof(T) == 8 ? 4 : 1;
  const bool groups = HH > 1 && ke - kb <= kMaxPass * kPerPass;
  const int32_t le = (int32_t)threadIdx.x / L, l = (int32_t)threadIdx.x - le * L;
  int32_t ra[kMaxPass], rz[kMaxPass];
  int32_t d0[kMaxPass], d1[kMaxPass];  // entry ids (< 2^31), -1: none
#pragma unroll
  for (int ps = 0; ps < kMaxPass; ++ps) {
    const int32_t k = kb + le + ps * kPerPass;
    const bool has = groups && le < kPerPass && k < ke;
    ra[ps] = has ? seg_ptr[k] - cb : 0;
    rz[ps] = has ? seg_ptr[k + 1] - cb : 0;
    d0[ps] = !has ? -1 : entry ? entry[k] : k;
    d1[ps] = (has && entry) ? entry_t[k] : -1;
  }
  if (!seg_slot) {
    // contiguous contributions: the run's slots are one contiguous range,
    // read as 16-byte vectors from the vector boundary at or before it (the
    // leading values belong to the previous run); a vector reaching past the
    // run's end is read value by value, so nothing past the last run is
    // touched.  Every vector of this thread is loaded before any value is
    // stored.  Each slot is read once: non-temporal loads leave the caches to
    // the output entries (the scattered mirror writes merge there), measured
    // 2-5 % faster on the 3 x 3 block reductions.
    constexpr int kW = 16 / (int)sizeof(T);  // values per vector
    typedef T vec_t __attribute__((ext_vector_type(kW)));
    const int64_t base = (int64_t)cb * HH;
    const int32_t lead = (int32_t)(base & (kW - 1));
    const T* src = slot + base;
    const vec_t* vsrc = reinterpret_cast<const vec_t*>(slot + (base - lead));
    const int32_t nv = n * HH;
    const int32_t nvec = (lead + nv + kW - 1) / kW;
    constexpr int kPerThread = ((kCap * HH + 2 * (kW - 1)) / kW + kBlock - 1) / kBlock;
    vec_t v[kPerThread];
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) {
      const int32_t q = (int32_t)threadIdx.x + j * kBlock;
      const int32_t i0 = q * kW - lead;  // run index of the vector's first value
      if (q < nvec && i0 + kW <= nv) {
        v[j] = __builtin_nontemporal_load(&vsrc[q]);
      } else {
#pragma unroll
        for (int e = 0; e < kW; ++e) v[j][e] = (q < nvec && i0 + e >= 0 && i0 + e < nv) ? src[i0 + e] : T(0);
      }
    }
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) {
      const int32_t q = (int32_t)threadIdx.x + j * kBlock;
      const int32_t i0 = q * kW - lead;
#pragma unroll
      for (int e = 0; e < kW; ++e)
        if (q < nvec && i0 + e >= 0 && i0 + e < nv) lds[i0 + e] = v[j][e];
    }
  } else if constexpr (HH == 1) {
    int32_t t = threadIdx.x;
    // four independent gathers per iteration keep the memory pipeline full
    for (; t + 3 * kBlock < n; t += 4 * kBlock) {
      const int32_t s0 = seg_slot[cb + t] & INT32_MAX, s1 = seg_slot[cb + t + kBlock] & INT32_MAX;
      const int32_t s2 = seg_slot[cb + t + 2 * kBlock] & INT32_MAX, s3 = seg_slot[cb + t + 3 * kBlock] & INT32_MAX;
      const T v0 = slot[s0], v1 = slot[s1], v2 = slot[s2], v3 = slot[s3];
      lds[t] = v0;
      lds[t + kBlock] = v1;
      lds[t + 2 * kBlock] = v2;
      lds[t + 3 * kBlock] = v3;
    }
    for (; t < n; t += kBlock) lds[t] = slot[seg_slot[cb + t] & INT32_MAX];
  } else {
    // the run's slot ids go to LDS in one coalesced pass, so that each value
    // gather below is a single global round trip; consecutive threads then
    // read consecutive values of a slot, kGather independent loads in flight
    __shared__ int32_t s_ids[kCap];
    for (int32_t t = threadIdx.x; t < n; t += kBlock) s_ids[t] = seg_slot[cb + t];
    __syncthreads();
    constexpr int kGather = 4;
    const int32_t nv = n * HH;
    int32_t u = threadIdx.x;
    for (; u + (kGather - 1) * kBlock < nv; u += kGather * kBlock) {
      int32_t t[kGather], h[kGather];
#pragma unroll
      for (int k = 0; k < kGather; ++k) {
        t[k] = (u + k * kBlock) / HH;
        h[k] = u + k * kBlock - t[k] * HH;
      }
      T v[kGather];
#pragma unroll
      for (int k = 0; k < kGather; ++k) v[k] = slot[slot_value<HH>(s_ids[t[k]], h[k])];
#pragma unroll
      for (int k = 0; k < kGather; ++k) lds[u + k * kBlock] = v[k];
    }
    for (; u < nv; u += kBlock) {
      const int32_t t = u / HH, h = u - t * HH;
      lds[u] = slot[slot_value<HH>(s_ids[t], h)];
    }
  }
  __syncthreads();
  if (!groups) {
    for (int32_t k = kb + threadIdx.x; k < ke; k += kBlock) {
      const int32_t a = seg_ptr[k] - cb, z = seg_ptr[k + 1] - cb;
      T acc[HH];
#pragma unroll
      for (int h = 0; h < HH; ++h) acc[h] = T(0);
      for (int32_t u = a; u < z; ++u) {
#pragma unroll
        for (int h = 0; h < HH; ++h) acc[h] += lds[u * HH + h];
      }
      write_entry<T, HH>(out, k, acc, entry, entry_t);
    }
  } else {
    // L lanes per entry (one per block row for D x D blocks), lane l summing
    // values [l*V, l*V + V), up to kMaxPass entries per lane group.  The
    // lanes of an entry read adjacent LDS words and write its HH values (and
    // their transposed copy) as one contiguous piece, where a thread per
    // entry stores HH-strided words from every lane.  Each value is still
    // summed over the entry's contributions in order.
    constexpr int V = HH / L;
#pragma unroll
    for (int ps = 0; ps < kMaxPass; ++ps) {
      if (d0[ps] < 0) break;
      T acc[V];
#pragma unroll
      for (int v = 0; v < V; ++v) acc[v] = T(0);
      for (int32_t u = ra[ps]; u < rz[ps]; ++u) {
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v] += lds[u * HH + l * V + v];
      }
#pragma unroll
      for (int v = 0; v < V; ++v) out[(int64_t)d0[ps] * HH + l * V + v] = acc[v];
      if (d1[ps] >= 0 && d1[ps] != d0[ps]) {
        // value (l, v) of the block lands at (v, l) of the mirror
#pragma unroll
        for (int v = 0; v < V; ++v) out[(int64_t)d1[ps] * HH + (kD > 1 ? v * kD + l : l * V + v)] = acc[v];
      }
    }
  }
}

template <typename T, int HH>
int launch_hh(int64_t nnz, const int32_t* seg_ptr, const int32_t* seg_slot, const int32_t* entry,
              const int32_t* entry_t, int64_t n_blocks, const int32_t* blk_ptr, const void* slot, void* out,
              hipStream_t stream) {
  const T* s = static_cast<const T*>(slot);
  T* o = static_cast<T*>(out);
  if (n_blocks > 0 && blk_ptr) {
    reduce_stream_kernel<T, HH><<<n_blocks, kBlock, 0, stream>>>(blk_ptr, n_blocks, seg_ptr, seg_slot, s, entry,
                                                                  entry_t, o);
  } else {
    const int64_t grid = (nnz + kBlock - 1) / kBlock;
    reduce_walk_kernel<T, HH><<<grid, kBlock, 0, stream>>>(seg_ptr, seg_slot, s, nnz, entry, entry_t, o);
  }
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("segmented reduce launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
}

template <typename T>
int launch(int64_t nnz, int hh, const int32_t* seg_ptr, const int32_t* seg_slot, const int32_t* entry,
           const int32_t* entry_t, int64_t n_blocks, const int32_t* blk_ptr, const void* slot, void* out,
           hipStream_t stream) {
#define TM_HH(H) launch_hh<T, H>(nnz, seg_ptr, seg_slot, entry, entry_t, n_blocks, blk_ptr, slot, out, stream)
  switch (hh) {
    case 1: return TM_HH(1);
    case 2: return TM_HH(2);
    case 3: return TM_HH(3);
    case 4: return TM_HH(4);
    case 9: return TM_HH(9);
    default: return tmh::fail(TM_ERR_UNSUPPORTED, "values per slot must be 1, 2, 3, 4 or 9");
  }
#undef TM_HH
}

// contiguous: contribution t is slot t (seg_slot is not read and may be NULL)
int reduce_checked(int32_t dtype, int64_t n_seg, int32_t hh, const int32_t* seg_ptr, const int32_t* seg_slot,
                   const int32_t* entry, const int32_t* entry_t, int64_t n_blocks, const int32_t* blk_ptr,
                   const void* slot, void* out, hipStream_t stream, bool contiguous = false) {
  if (n_seg < 0 || n_seg >= tmh::kMaxIndex || n_blocks < 0 || n_blocks > n_seg)
    return tmh::fail(TM_ERR_SIZE, "nnz / plan size out of range");
  if (n_seg == 0) return TM_OK;
  if (contiguous) seg_slot = nullptr;
  if (!seg_ptr || (!seg_slot && !contiguous) || !slot || !out) return tmh::fail(TM_ERR_ARG, "NULL pointer");
  if ((entry == nullptr) != (entry_t == nullptr)) return tmh::fail(TM_ERR_ARG, "entry and entry_t go together");
  if (dtype == TM_F64)
    return launch<double>(n_seg, hh, seg_ptr, seg_slot, entry, entry_t, n_blocks, blk_ptr, slot, out, stream);
  if (dtype == TM_F32)
    return launch<float>(n_seg, hh, seg_ptr, seg_slot, entry, entry_t, n_blocks, blk_ptr, slot, out, stream);
  return tmh::fail(TM_ERR_ARG, "unknown dtype");
}

// Slot s = e*B*B + i*B + j becomes the packed upper slot e*P + p(a, b) of
// (a, b) = (min(i, j), max(i, j)), P = B(B+1)/2, p(a, b) = a*B - a(a-1)/2 + b - a
// (row-major over a <= b), with bit 31 set when i > j (read transposed).
__global__ void __launch_bounds__(kBlock) symmetric_slots_kernel(const int32_t* __restrict__ seg_slot, int64_t n,
                                                                 int32_t B, int32_t* __restrict__ sym) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n) return;
  const int32_t s = seg_slot[t];
  const int32_t BB = B * B, P = B * (B + 1) / 2;
  const int32_t e = s / BB, r = s % BB, i = r / B, j = r % B;
  const int32_t a = i < j ? i : j, b = i < j ? j : i;
  const int32_t packed = e * P + a * B - a * (a - 1) / 2 + (b - a);
  sym[t] = i <= j ? packed : (packed | kTransposed);
}

}  // namespace

extern "C" int tm_symmetric_segments(int32_t n_basis, int64_t n_slots, const int32_t* seg_slot, int32_t* sym_slot,
                                     tm_stream_t stream_) {
  TM_TRY_BEGIN
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  if (n_basis < 1 || n_basis > 255) return tmh::fail(TM_ERR_ARG, "n_basis out of range");
  if (n_slots < 0 || n_slots >= tmh::kMaxIndex || n_slots % ((int64_t)n_basis * n_basis) != 0)
    return tmh::fail(TM_ERR_SIZE, "n_slots must be n_elem * n_basis^2 < 2^31");
  if (n_slots == 0) return TM_OK;
  if (!seg_slot || !sym_slot) return tmh::fail(TM_ERR_ARG, "NULL pointer");
  const int64_t grid = (n_slots + kBlock - 1) / kBlock;
  symmetric_slots_kernel<<<grid, kBlock, 0, stream>>>(seg_slot, n_slots, n_basis, sym_slot);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("symmetric segments launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
  TM_TRY_END
}

extern "C" int tm_segmented_reduce(int32_t dtype, int64_t nnz, int32_t hh, const int32_t* seg_ptr,
                                   const int32_t* seg_slot, int64_t n_blocks, const int32_t* blk_ptr,
                                   const void* slot, void* out, tm_stream_t stream_) {
  TM_TRY_BEGIN
  return reduce_checked(dtype, nnz, hh, seg_ptr, seg_slot, nullptr, nullptr, n_blocks, blk_ptr, slot, out,
                        reinterpret_cast<hipStream_t>(stream_));
  TM_TRY_END
}

extern "C" int tm_segmented_reduce_pairs(int32_t dtype, int64_t n_seg, int32_t hh, const int32_t* seg_ptr,
                                         const int32_t* seg_slot, const int32_t* entry, const int32_t* entry_t,
                                         int64_t n_blocks, const int32_t* blk_ptr, const void* slot, void* out,
                                         tm_stream_t stream_) {
  TM_TRY_BEGIN
  if (n_seg > 0 && (!entry || !entry_t)) return tmh::fail(TM_ERR_ARG, "entry / entry_t is NULL");
  return reduce_checked(dtype, n_seg, hh, seg_ptr, seg_slot, entry, entry_t, n_blocks, blk_ptr, slot, out,
                        reinterpret_cast<hipStream_t>(stream_));
  TM_TRY_END
}

int tmh::reduce_view(int32_t dtype, const tm_pattern_view* pat, int32_t hh, const void* slot, void* out,
                     tm_stream_t stream) {
  return reduce_checked(dtype, pat->nnz, hh, pat->seg_ptr, pat->seg_slot, pat->entry, pat->entry_t, pat->n_blocks,
                        pat->blk_ptr, slot, out, reinterpret_cast<hipStream_t>(stream), pat->slot_dest != nullptr);
}

namespace {
__global__ void __launch_bounds__(kBlock) slot_dest_kernel(const int32_t* __restrict__ seg_slot, int64_t n,
                                                           int32_t* __restrict__ dest) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n) return;
  const int32_t raw = seg_slot[t];
  dest[raw & INT32_MAX] = (int32_t)t | (raw & kTransposed);
}
}  // namespace

extern "C" int tm_slot_dest_build(int64_t n, const int32_t* seg_slot, int32_t* slot_dest, tm_stream_t stream_) {
  TM_TRY_BEGIN
  if (n < 0 || n >= tmh::kMaxIndex) return tmh::fail(TM_ERR_SIZE, "n out of range");
  if (n == 0) return TM_OK;
  if (!seg_slot || !slot_dest) return tmh::fail(TM_ERR_ARG, "NULL pointer");
  const int64_t grid = (n + kBlock - 1) / kBlock;
  slot_dest_kernel<<<grid, kBlock, 0, reinterpret_cast<hipStream_t>(stream_)>>>(seg_slot, n, slot_dest);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("slot destinations launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
  TM_TRY_END
}
