// Linear-form (load vector) assembly.
//
// Replaces NodeAssembler.__call__ for the load integrands f * v (nodal f,
// README.md:103-118) and c * v (builtin.py:730-770):
//   tensormesh/assemble/node_assembler.py:357-409
// The nodal field is interpolated to the quadrature points with the shape
// functions (f_h(q) = sum_b f[n_b] phi_b(q), node_assembler.py:388-390), the
// integrand f_h(q) phi_i(q) is integrated against jxw_q = w_q |det J_q| one
// quadrature point after another (the default batch_size=1 order), and the
// per-element vectors are summed into nodes.  The scatter of the reference
// (ReduceProjector over conn.flatten(), node_assembler.py:665-670) becomes a
// gather over the node incidence lists (ascending e*B + i), done by the
// segmented reduction of tm_reduce.hip with seg_ptr = inc_ptr, seg_slot = inc.
//
// tm_assemble_linear_noalloc runs the same two stages with the slot values in
// a file-scope device array (`scratch`) instead of a caller workspace, for
// callers that cannot allocate.  Its device address is looked up once per GPU
// with hipGetSymbolAddress (the host-side name of a __device__ array is not
// a device address) and kept per device.
#include <hip/hip_runtime.h>

#include <mutex>

#include "tm_internal.h"

namespace {

constexpr int kBlock = 256;

// Slot values of tm_assemble_linear_noalloc: n_elem*B values of the mesh dtype.
static __device__ __align__(256) unsigned char scratch[TM_LINEAR_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call on that device (with it current) and kept.
int scratch_on_current_device(void** out) {
  int dev = -1;
  hipError_t err = hipGetDevice(&dev);
  if (err != hipSuccess) {
    tmh::set_error("hipGetDevice: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  if (dev < 0 || dev >= kMaxDevices) return tmh::fail(TM_ERR_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (err != hipSuccess || !p) {
      tmh::set_error("hipGetSymbolAddress(scratch): %s", hipGetErrorString(err));
      return TM_ERR_HIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return TM_OK;
}

template <typename T, int D, int B, bool CONST>
__global__ void __launch_bounds__(kBlock)
    element_load_kernel(const int32_t* __restrict__ conn, const T* __restrict__ pts, const T* __restrict__ f,
                        double c, const double* __restrict__ tab, int Q, int64_t n_elem, T* __restrict__ slot) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= n_elem) return;
  const double* w = tab;
  const double* phi = tab + Q;
  const double* dphi = tab + Q + Q * B;
  int v[B];
  double x[B][D];
  double fb[B];
#pragma unroll
  for (int b = 0; b < B; ++b) {
    v[b] = conn[e * B + b];
#pragma unroll
    for (int d = 0; d < D; ++d) x[b][d] = (double)pts[(int64_t)v[b] * D + d];
    fb[b] = CONST ? c : (double)f[v[b]];
  }
  double acc[B];
#pragma unroll
  for (int i = 0; i < B; ++i) acc[i] = 0.0;
  for (int q = 0; q < Q; ++q) {
    double J[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < B; ++b) s += x[b][j] * dphi[(q * D + i) * B + b];
        J[i][j] = s;
      }
    double det;
    if constexpr (D == 2) {
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    } else {
      det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) + J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
            J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    }
    const double jxw = w[q] * fabs(det);
    double fq = 0.0;
#pragma unroll
    for (int b = 0; b < B; ++b) fq += fb[b] * phi[q * B + b];
#pragma unroll
    for (int i = 0; i < B; ++i) acc[i] += (fq * phi[q * B + i]) * jxw;
  }
  T* out = slot + e * B;
#pragma unroll
  for (int i = 0; i < B; ++i) out[i] = (T)acc[i];
}

template <typename T, int D, int B>
int launch(const tm_mesh_view* m, const tm_form_desc* f, const double* cval, void* slot_ws, hipStream_t s) {
  const int64_t grid = (m->n_elem + kBlock - 1) / kBlock;
  if (grid == 0) return TM_OK;
  const T* pts = static_cast<const T*>(m->points);
  const T* fv = static_cast<const T*>(f->coeff);
  T* slot = static_cast<T*>(slot_ws);
  if (cval)
    element_load_kernel<T, D, B, true><<<grid, kBlock, 0, s>>>(m->conn, pts, fv, *cval, f->tab_dev, f->n_quad, m->n_elem, slot);
  else
    element_load_kernel<T, D, B, false><<<grid, kBlock, 0, s>>>(m->conn, pts, fv, 0.0, f->tab_dev, f->n_quad, m->n_elem, slot);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    tmh::set_error("load kernel launch: %s", hipGetErrorString(err));
    return TM_ERR_HIP;
  }
  return TM_OK;
}

template <typename T>
int dispatch(const tm_mesh_view* m, const tm_form_desc* f, const double* cval, void* slot_ws, hipStream_t s) {
  switch (m->cell) {
#define TM_CASE(C, D, B) \
  case C: return launch<T, D, B>(m, f, cval, slot_ws, s);
    TM_ALL_CELLS(TM_CASE)
#undef TM_CASE
    default: return tmh::fail(TM_ERR_ARG, "unknown cell");
  }
}

int check_linear_args(const tm_mesh_view* mesh, const tm_form_desc* form, const double* const_value,
                      const tm_incidence_view* inc) {
  if (!mesh || !form || !inc) return tmh::fail(TM_ERR_ARG, "NULL argument");
  if (tmh::cell_basis(mesh->cell) < 0) return tmh::fail(TM_ERR_ARG, "unknown cell");
  if (mesh->dtype != TM_F64 && mesh->dtype != TM_F32) return tmh::fail(TM_ERR_ARG, "unknown dtype");
  if (mesh->n_elem < 0 || mesh->n_elem * tmh::cell_basis(mesh->cell) >= tmh::kMaxIndex)
    return tmh::fail(TM_ERR_SIZE, "mesh too large for int32 indexing");
  if (!form->tab_dev || form->n_quad <= 0) return tmh::fail(TM_ERR_ARG, "device form tables missing");
  if (!const_value && !form->coeff) return tmh::fail(TM_ERR_ARG, "need a nodal field or a constant");
  if (mesh->n_points < 0 || mesh->n_points >= tmh::kMaxIndex) return tmh::fail(TM_ERR_SIZE, "n_points out of range");
  if (mesh->n_points > 0 && (!inc->inc_ptr || !inc->inc)) return tmh::fail(TM_ERR_ARG, "NULL incidence lists");
  return TM_OK;
}

int run_linear(const tm_mesh_view* mesh, const tm_form_desc* form, const double* const_value,
               const tm_incidence_view* inc, void* slot_ws, void* out, tm_stream_t stream_) {
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  int st = mesh->dtype == TM_F64 ? dispatch<double>(mesh, form, const_value, slot_ws, stream)
                                 : dispatch<float>(mesh, form, const_value, slot_ws, stream);
  if (st) return st;
  return tm_segmented_reduce(mesh->dtype, mesh->n_points, 1, inc->inc_ptr, inc->inc, inc->n_blocks, inc->blk_ptr,
                             slot_ws, out, stream_);
}

}  // namespace

extern "C" int tm_assemble_linear(const tm_mesh_view* mesh, const tm_form_desc* form, const double* const_value,
                                  const tm_incidence_view* inc, void* slot_ws, void* out, tm_stream_t stream_) {
  TM_TRY_BEGIN
  int st = check_linear_args(mesh, form, const_value, inc);
  if (st) return st;
  if (mesh->n_elem > 0 && (!slot_ws || !mesh->conn || !mesh->points)) return tmh::fail(TM_ERR_ARG, "NULL buffer");
  if (mesh->n_points > 0 && !out) return tmh::fail(TM_ERR_ARG, "NULL output");
  return run_linear(mesh, form, const_value, inc, slot_ws, out, stream_);
  TM_TRY_END
}

extern "C" int tm_assemble_linear_noalloc(const tm_mesh_view* mesh, const tm_form_desc* form,
                                          const double* const_value, const tm_incidence_view* inc, void* out,
                                          tm_stream_t stream_) {
  TM_TRY_BEGIN
  int st = check_linear_args(mesh, form, const_value, inc);
  if (st) return st;
  const int64_t value_bytes = mesh->dtype == TM_F64 ? 8 : 4;
  const int64_t need = mesh->n_elem * tmh::cell_basis(mesh->cell) * value_bytes;
  if (need > (int64_t)TM_LINEAR_SCRATCH_BYTES) {
    tmh::set_error("tm_assemble_linear_noalloc: %lld slot bytes exceed the %lld-byte scratch array "
                   "(use tm_assemble_linear with a workspace)",
                   (long long)need, (long long)TM_LINEAR_SCRATCH_BYTES);
    return TM_ERR_SIZE;
  }
  if (mesh->n_elem > 0 && (!mesh->conn || !mesh->points)) return tmh::fail(TM_ERR_ARG, "NULL buffer");
  if (mesh->n_points > 0 && !out) return tmh::fail(TM_ERR_ARG, "NULL output");
  if (stream_) {
    // the scratch address is per device: the stream must belong to the current one
    int cur = -1, sdev = -1;
    if (hipGetDevice(&cur) != hipSuccess ||
        hipStreamGetDevice(reinterpret_cast<hipStream_t>(stream_), &sdev) != hipSuccess)
      return tmh::fail(TM_ERR_HIP, "tm_assemble_linear_noalloc: device query failed");
    if (cur != sdev) return tmh::fail(TM_ERR_ARG, "tm_assemble_linear_noalloc: stream is not on the current device");
  }
  void* slot = nullptr;
  st = scratch_on_current_device(&slot);
  if (st) return st;
  return run_linear(mesh, form, const_value, inc, slot, out, stream_);
  TM_TRY_END
}
