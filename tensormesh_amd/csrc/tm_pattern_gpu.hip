// Device-side pattern builder: the same arrays as tm_pattern_build /
// tm_incidence_build / tm_colpos_build, computed on the GPU from device
// connectivity.
//
// Replaces the edge / projector construction of ElementAssembler.from_mesh
// (tensormesh/assemble/element_assembler.py:804-853), whose scipy
// coo->csr->coo sort runs on the host after a device-to-host copy.  Steps:
//   1. node degrees with integer atomics (exact counts), exclusive scan ->
//      inc_ptr (one workgroup, chunked);
//   2. incidence fill with atomic cursors, then each node's list sorted by
//      e*B + i (insertion sort per node), so the result does not depend on
//      the atomic order;
//   3. one wavefront per CSR row: gather the neighbour ids conn[e, j] of the
//      row's incidences into LDS, bitonic sort, count the distinct ones
//      (row lengths), scan -> row_ptr;
//   4. one wavefront per row again: write the distinct columns, the packed
//      positions of every (incidence, j), per-entry contribution counts and
//      their scan -> seg_ptr, then every entry's slots in ascending element
//      order -> seg_slot, slot2nnz, entry masks.
// Workgroups are a single wavefront, so every barrier is a wave barrier and
// rows with different key counts never diverge across a workgroup barrier.
#include <hip/hip_runtime.h>

#include "tm_internal.h"

namespace {

constexpr int kWave = 64;
constexpr int kMaxKeys = 2048;  // neighbour keys (incidences x B) per row
constexpr int kScanBlock = 1024;

__global__ void __launch_bounds__(256) fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(256) count_kernel(const int32_t* __restrict__ conn, int64_t total,
                                                    int32_t* __restrict__ deg) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < total) atomicAdd(&deg[conn[t]], 1);
}

// out[0] = 0, out[i+1] = sum_{j<=i} in[i]; one workgroup walks the array in chunks.
__global__ void __launch_bounds__(kScanBlock) scan_kernel(const int32_t* __restrict__ in, int64_t n,
                                                          int32_t* __restrict__ out) {
  __shared__ int64_t sh[kScanBlock];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) {
    carry = 0;
    out[0] = 0;
  }
  __syncthreads();
  for (int64_t base = 0; base < n; base += kScanBlock) {
    const int64_t i = base + threadIdx.x;
    sh[threadIdx.x] = i < n ? in[i] : 0;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
      const int64_t v = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += v;
      __syncthreads();
    }
    if (i < n) out[i + 1] = (int32_t)(carry + sh[threadIdx.x]);
    __syncthreads();
    if (threadIdx.x == 0) carry += sh[kScanBlock - 1];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256) fill_inc_kernel(const int32_t* __restrict__ conn, int64_t total,
                                                       int32_t* __restrict__ cursor, int32_t* __restrict__ inc) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < total) inc[atomicAdd(&cursor[conn[t]], 1)] = (int32_t)t;
}

__global__ void __launch_bounds__(256) sort_inc_kernel(const int32_t* __restrict__ inc_ptr, int64_t n_points,
                                                       int32_t* __restrict__ inc) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= n_points) return;
  const int32_t a = inc_ptr[r], b = inc_ptr[r + 1];
  for (int32_t i = a + 1; i < b; ++i) {
    const int32_t v = inc[i];
    int32_t j = i - 1;
    while (j >= a && inc[j] > v) {
      inc[j + 1] = inc[j];
      --j;
    }
    inc[j + 1] = v;
  }
}

// Gather and sort the row's neighbour keys in LDS; returns the key count
// (sorted ascending in keys[0, n)), or -1 when the row exceeds kMaxKeys.
__device__ int gather_sorted(const int32_t* __restrict__ conn, int B, const int32_t* __restrict__ inc_ptr,
                             const int32_t* __restrict__ inc, int64_t r, int32_t* keys) {
  const int32_t a = inc_ptr[r], b = inc_ptr[r + 1];
  const int n = (b - a) * B;
  if (n > kMaxKeys) return -1;
  int m = 1;
  while (m < n) m <<= 1;
  for (int k = threadIdx.x; k < m; k += kWave) {
    if (k < n) {
      const int32_t ei = inc[a + k / B];
      keys[k] = conn[(int64_t)(ei / B) * B + (k % B)];
    } else {
      keys[k] = 0x7fffffff;
    }
  }
  __syncthreads();
  for (int size = 2; size <= m; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int k = threadIdx.x; k < m; k += kWave) {
        const int partner = k ^ stride;
        if (partner > k) {
          const bool up = (k & size) == 0;
          const int32_t x = keys[k], y = keys[partner];
          if ((x > y) == up) {
            keys[k] = y;
            keys[partner] = x;
          }
        }
      }
      __syncthreads();
    }
  }
  return n;
}

__global__ void __launch_bounds__(kWave) row_count_kernel(const int32_t* __restrict__ conn, int B,
                                                          const int32_t* __restrict__ inc_ptr,
                                                          const int32_t* __restrict__ inc, int64_t n_points,
                                                          int32_t* __restrict__ rowlen, int32_t* __restrict__ flag) {
  __shared__ int32_t keys[kMaxKeys];
  const int64_t r = blockIdx.x;
  if (r >= n_points) return;
  const int n = gather_sorted(conn, B, inc_ptr, inc, r, keys);
  if (n < 0) {
    if (threadIdx.x == 0) {
      rowlen[r] = 0;
      atomicOr(flag, 1);
    }
    return;
  }
  int cnt = 0;
  for (int k = threadIdx.x; k < n; k += kWave) cnt += (k == 0 || keys[k] != keys[k - 1]) ? 1 : 0;
  // wave sum through LDS (reuse the tail of keys is unsafe: use a small array)
  __shared__ int32_t part[kWave];
  part[threadIdx.x] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int l = 0; l < kWave; ++l) s += part[l];
    rowlen[r] = s;
  }
}

__global__ void __launch_bounds__(kWave)
    row_fill_kernel(const int32_t* __restrict__ conn, int B, const int32_t* __restrict__ inc_ptr,
                    const int32_t* __restrict__ inc, const int32_t* __restrict__ row_ptr, int64_t n_points,
                    int32_t* __restrict__ col_idx, int32_t* __restrict__ seg_ptr, int32_t* __restrict__ seg_slot,
                    int32_t* __restrict__ slot2nnz, uint32_t* __restrict__ colpos, uint32_t* __restrict__ entry_mask) {
  __shared__ int32_t keys[kMaxKeys];
  __shared__ int32_t uniq[kMaxKeys];
  __shared__ int16_t pos[kMaxKeys];      // position of key (t, j) in the row (unsorted order)
  __shared__ int32_t cnt[kMaxKeys];      // contributions per entry, then their scan
  const int64_t r = blockIdx.x;
  if (r >= n_points) return;
  const int n = gather_sorted(conn, B, inc_ptr, inc, r, keys);
  if (n < 0) return;  // flagged by row_count_kernel
  const int32_t k0 = row_ptr[r], len = row_ptr[r + 1] - row_ptr[r];
  // distinct columns, in order: one lane walks (rows are short)
  if (threadIdx.x == 0) {
    int u = 0;
    for (int k = 0; k < n; ++k)
      if (k == 0 || keys[k] != keys[k - 1]) uniq[u++] = keys[k];
  }
  __syncthreads();
  for (int p = threadIdx.x; p < len; p += kWave) {
    col_idx[k0 + p] = uniq[p];
    cnt[p] = 0;
  }
  const int32_t a = inc_ptr[r];
  const int n_inc = inc_ptr[r + 1] - a;
  // position of every (incidence t, local node j)
  for (int k = threadIdx.x; k < n; k += kWave) {
    const int32_t ei = inc[a + k / B];
    const int32_t c = conn[(int64_t)(ei / B) * B + (k % B)];
    int lo = 0, hi = len - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (uniq[mid] < c) lo = mid + 1; else hi = mid;
    }
    pos[k] = (int16_t)lo;
  }
  __syncthreads();
  if (colpos && B <= 4) {
    for (int t = threadIdx.x; t < n_inc; t += kWave) {
      uint32_t w = 0;
      for (int j = 0; j < B; ++j) w |= (uint32_t)pos[t * B + j] << (8 * j);
      colpos[a + t] = w;
    }
  }
  // contributions per entry, then their exclusive scan (one lane: rows are short)
  if (threadIdx.x == 0) {
    for (int k = 0; k < n; ++k) cnt[pos[k]] += 1;
    int32_t run = 0;
    for (int p = 0; p < len; ++p) {
      const int32_t c = cnt[p];
      cnt[p] = run;
      run += c;
    }
  }
  __syncthreads();
  const int64_t sbase = (int64_t)a * B;  // this row's first contribution
  for (int p = threadIdx.x; p < len; p += kWave) {
    seg_ptr[k0 + p] = (int32_t)(sbase + cnt[p]);
    int32_t w = (int32_t)(sbase + cnt[p]);
    uint32_t mask = 0;
    for (int t = 0; t < n_inc; ++t) {
      const int32_t ei = inc[a + t];
      const int64_t e = ei / B, i = ei % B;
      for (int j = 0; j < B; ++j) {
        if (pos[t * B + j] == p) {
          const int32_t s = (int32_t)(e * B * B + i * B + j);
          seg_slot[w++] = s;
          slot2nnz[s] = k0 + p;
          if (t < 32) mask |= 1u << t;
        }
      }
    }
    if (entry_mask) entry_mask[k0 + p] = n_inc <= 32 ? mask : 0u;
  }
}

__global__ void copy_i32(const int32_t* __restrict__ src, int64_t n, int32_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ void seg_end_kernel(const int32_t* __restrict__ row_ptr, int64_t n_points, int32_t* __restrict__ seg_ptr,
                               int32_t total_slots) {
  if (threadIdx.x == 0 && blockIdx.x == 0) seg_ptr[row_ptr[n_points]] = total_slots;
}

int hip_ok(const char* what) {
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("%s: %s", what, hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
}

}  // namespace

extern "C" int tm_pattern_device_count(const int32_t* conn, int64_t n_elem, int32_t n_basis, int64_t n_points,
                                       int32_t* inc_ptr, int32_t* inc, int32_t* row_ptr, int32_t* work,
                                       tm_stream_t stream_) {
  TM_TRY_BEGIN
  hipStream_t s = reinterpret_cast<hipStream_t>(stream_);
  if (!conn || !inc_ptr || !inc || !row_ptr || !work) return tmh::fail(TM_ERR_ARG, "NULL argument");
  if (n_elem < 0 || n_basis <= 0 || n_basis > 27 || n_points <= 0) return tmh::fail(TM_ERR_ARG, "bad sizes");
  if (n_elem * n_basis * n_basis >= tmh::kMaxIndex || n_points >= tmh::kMaxIndex)
    return tmh::fail(TM_ERR_SIZE, "pattern exceeds int32 indexing");
  const int64_t total = n_elem * n_basis;
  // work: [n_points] scratch (degrees, cursors, row lengths) + [1] overflow flag
  int32_t* deg = work;
  int32_t* flag = work + n_points;
  const int64_t gp = (n_points + 1 + 255) / 256, gt = (total + 255) / 256;
  fill_i32<<<gp, 256, 0, s>>>(work, n_points + 1, 0);
  if (total > 0) count_kernel<<<gt, 256, 0, s>>>(conn, total, deg);
  scan_kernel<<<1, kScanBlock, 0, s>>>(deg, n_points, inc_ptr);
  // cursors = inc_ptr[0, n_points)
  copy_i32<<<(n_points + 255) / 256, 256, 0, s>>>(inc_ptr, n_points, deg);
  if (total > 0) fill_inc_kernel<<<gt, 256, 0, s>>>(conn, total, deg, inc);
  sort_inc_kernel<<<(n_points + 255) / 256, 256, 0, s>>>(inc_ptr, n_points, inc);
  row_count_kernel<<<n_points, kWave, 0, s>>>(conn, n_basis, inc_ptr, inc, n_points, deg, flag);
  scan_kernel<<<1, kScanBlock, 0, s>>>(deg, n_points, row_ptr);
  return hip_ok("device pattern count");
  TM_TRY_END
}

extern "C" int tm_pattern_device_fill(const int32_t* conn, int64_t n_elem, int32_t n_basis, int64_t n_points,
                                      const int32_t* inc_ptr, const int32_t* inc, const int32_t* row_ptr,
                                      int32_t* col_idx, int32_t* seg_ptr, int32_t* seg_slot, int32_t* slot2nnz,
                                      uint32_t* colpos, uint32_t* entry_mask, tm_stream_t stream_) {
  TM_TRY_BEGIN
  hipStream_t s = reinterpret_cast<hipStream_t>(stream_);
  if (!conn || !inc_ptr || !inc || !row_ptr || !col_idx || !seg_ptr || !seg_slot || !slot2nnz)
    return tmh::fail(TM_ERR_ARG, "NULL argument");
  if (n_basis <= 0 || n_basis > 27 || n_points <= 0) return tmh::fail(TM_ERR_ARG, "bad sizes");
  row_fill_kernel<<<n_points, kWave, 0, s>>>(conn, n_basis, inc_ptr, inc, row_ptr, n_points, col_idx, seg_ptr,
                                             seg_slot, slot2nnz, n_basis <= 4 ? colpos : nullptr,
                                             n_basis <= 4 ? entry_mask : nullptr);
  seg_end_kernel<<<1, 64, 0, s>>>(row_ptr, n_points, seg_ptr, (int32_t)(n_elem * n_basis * n_basis));
  return hip_ok("device pattern fill");
  TM_TRY_END
}
