// Error reporting and version query of the C-ABI.
#include <cstdarg>
#include <cstdio>

#include "tm_internal.h"

namespace {
thread_local char g_last_error[512] = "";
}

namespace tmh {
void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}
}  // namespace tmh

extern "C" int tm_abi_version(void) { return TM_ABI_VERSION; }

extern "C" const char* tm_last_error(void) { return g_last_error; }

extern "C" int64_t tm_workspace_bytes(int32_t op, const tm_mesh_view* mesh, int32_t hh) {
  if (!mesh) return tmh::fail(TM_ERR_ARG, "mesh is NULL");
  const int64_t B = tmh::cell_basis(mesh->cell);
  if (B < 0) return tmh::fail(TM_ERR_ARG, "unknown cell");
  if (mesh->dtype != TM_F64 && mesh->dtype != TM_F32) return tmh::fail(TM_ERR_ARG, "unknown dtype");
  if (mesh->n_elem < 0 || mesh->n_elem >= tmh::kMaxIndex) return tmh::fail(TM_ERR_SIZE, "n_elem out of range");
  const int64_t s = mesh->dtype == TM_F64 ? 8 : 4;
  switch (op) {
    case TM_WS_ELEMENT_SLOTS:
      if (hh < 1 || hh > 9) return tmh::fail(TM_ERR_ARG, "hh must be in [1, 9]");
      return mesh->n_elem * B * B * hh * s;
    case TM_WS_NODE_SLOTS:
      if (hh < 1 || hh > 9) return tmh::fail(TM_ERR_ARG, "hh must be in [1, 9]");
      return mesh->n_elem * B * hh * s;
    case TM_WS_ENERGY:
      return (mesh->n_elem + TM_J2_ENERGY_PARTS) * 8;
    case TM_WS_ELEMENT_SLOTS_UPPER:
      if (hh < 1 || hh > 9) return tmh::fail(TM_ERR_ARG, "hh must be in [1, 9]");
      return mesh->n_elem * (B * (B + 1) / 2) * hh * s;
    default:
      return tmh::fail(TM_ERR_ARG, "unknown workspace");
  }
}
