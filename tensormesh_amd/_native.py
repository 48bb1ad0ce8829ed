"""ctypes binding of libtmhip.so (the C-ABI declared in include/tensormesh_amd.h).

The library is loaded from the package directory (built in-tree by
``tensormesh_amd.build``).  There is no fallback: if the library is missing
or does not export the expected ABI version, ``lib()`` raises.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path
from typing import Optional

import torch

LIB_PATH = Path(__file__).resolve().parent / "libtmhip.so"
ABI_VERSION = 3

# enum values (include/tensormesh_amd.h)
TM_F32, TM_F64 = 0, 1
CELL = {"triangle": 0, "tetra": 1, "triangle6": 2, "tetra10": 3,
        "quad": 4, "quad9": 5, "hexahedron": 6, "hexahedron27": 7, "triangle10": 8, "tetra20": 9}
FORM = {"laplace": 0, "diffusion_elem": 1, "diffusion_node": 2, "mass": 3, "linear_elasticity": 4}

c_i32p = ctypes.POINTER(ctypes.c_int32)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_dblp = ctypes.POINTER(ctypes.c_double)


class MeshView(ctypes.Structure):
    _fields_ = [("cell", ctypes.c_int32), ("dtype", ctypes.c_int32),
                ("n_elem", ctypes.c_int64), ("n_points", ctypes.c_int64),
                ("conn", ctypes.c_void_p), ("points", ctypes.c_void_p)]


class FormDesc(ctypes.Structure):
    _fields_ = [("form", ctypes.c_int32), ("n_quad", ctypes.c_int32),
                ("tab_host", ctypes.c_void_p), ("tab_dev", ctypes.c_void_p),
                ("coeff", ctypes.c_void_p), ("young", ctypes.c_double),
                ("poisson", ctypes.c_double), ("flags", ctypes.c_int32), ("epw", ctypes.c_int32)]


TM_FLAG_MFMA = 1
TM_FLAG_UPPER = 2   # symmetric forms: element blocks i <= j only, symmetric segment list
TM_FLAG_F32_NO_REDO = 4   # float32 P1 geometry for every element (no float64 redo of slivers)


class PatternView(ctypes.Structure):
    _fields_ = [("nnz", ctypes.c_int64), ("row_ptr", ctypes.c_void_p), ("col_idx", ctypes.c_void_p),
                ("seg_ptr", ctypes.c_void_p), ("seg_slot", ctypes.c_void_p),
                ("n_blocks", ctypes.c_int64), ("blk_ptr", ctypes.c_void_p),
                ("entry", ctypes.c_void_p), ("entry_t", ctypes.c_void_p), ("slot_dest", ctypes.c_void_p)]


class BlockRowsView(ctypes.Structure):
    _fields_ = [("nnz", ctypes.c_int64), ("row_ptr", ctypes.c_void_p), ("seg_ptr", ctypes.c_void_p),
                ("pairs", ctypes.c_void_p), ("inc_ptr", ctypes.c_void_p), ("inc", ctypes.c_void_p),
                ("n_blocks", ctypes.c_int64), ("blk_info", ctypes.c_void_p)]


class FacetView(ctypes.Structure):
    _fields_ = [("n_sel", ctypes.c_int64), ("elem", ctypes.c_void_p), ("local", ctypes.c_void_p)]


class RowsView(ctypes.Structure):
    _fields_ = [("nnz", ctypes.c_int64), ("row_ptr", ctypes.c_void_p), ("inc_ptr", ctypes.c_void_p),
                ("inc", ctypes.c_void_p), ("colpos", ctypes.c_void_p), ("entry_mask", ctypes.c_void_p),
                ("n_blocks", ctypes.c_int64),
                ("blk_ptr", ctypes.c_void_p), ("blk_info", ctypes.c_void_p),
                ("dest", ctypes.c_void_p), ("ent_start", ctypes.c_void_p)]


class PcgView(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int32), ("reserved", ctypes.c_int32), ("n_rows", ctypes.c_int64),
                ("n_sys", ctypes.c_int64), ("n_blocks", ctypes.c_int64)]
    _fields_ += [(n, ctypes.c_void_p) for n in ("row_ptr", "col_idx", "values", "dinv", "r", "z", "p", "q",
                                                 "partials", "scalars", "blk_sys", "blk_row0", "blk_row1",
                                                 "sys_blk")]
    _fields_ += [("rtol", ctypes.c_double), ("s", ctypes.c_void_p), ("two_launch", ctypes.c_int32),
                 ("parity", ctypes.c_int32), ("aux", ctypes.c_void_p)]


class J2Material(ctypes.Structure):
    _fields_ = [("young", ctypes.c_double), ("poisson", ctypes.c_double), ("sig0", ctypes.c_double),
                ("hardening", ctypes.c_double)]


class J2State(ctypes.Structure):
    _fields_ = [("eps_p", ctypes.c_void_p), ("alpha", ctypes.c_void_p)]


J2_ENERGY_PARTS = 1024
LINEAR_SCRATCH_BYTES = 64 << 20   # TM_LINEAR_SCRATCH_BYTES: slot array of tm_assemble_linear_noalloc


class ContactPlane(ctypes.Structure):
    _fields_ = [("kappa", ctypes.c_double), ("normal", ctypes.c_double * 3), ("offset", ctypes.c_double)]


class IncidenceView(ctypes.Structure):
    _fields_ = [("inc_ptr", ctypes.c_void_p), ("inc", ctypes.c_void_p),
                ("n_blocks", ctypes.c_int64), ("blk_ptr", ctypes.c_void_p)]


# name -> (restype, argtypes); every symbol declared in include/tensormesh_amd.h
SIGNATURES = {
    "tm_abi_version": (ctypes.c_int, []),
    "tm_workspace_bytes": (ctypes.c_int64, [ctypes.c_int32, ctypes.POINTER(MeshView), ctypes.c_int32]),
    "tm_last_error": (ctypes.c_char_p, []),
    "tm_pattern_nnz": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int64, c_i64p]),
    "tm_pattern_build": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int64,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p]),
    "tm_incidence_build": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int64,
                                          ctypes.c_void_p, ctypes.c_void_p]),
    "tm_assemble_bilinear": (ctypes.c_int, [ctypes.POINTER(MeshView), ctypes.POINTER(FormDesc),
                                            ctypes.POINTER(PatternView), ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p]),
    "tm_element_matrices": (ctypes.c_int, [ctypes.POINTER(MeshView), ctypes.POINTER(FormDesc),
                                           ctypes.c_void_p, ctypes.c_void_p]),
    "tm_segmented_reduce": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p]),
    "tm_reduce_plan": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p, c_i64p]),
    "tm_symmetric_segments": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p]),
    "tm_slot_dest_build": (ctypes.c_int, [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "tm_segmented_reduce_pairs": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_void_p]),
    "tm_assemble_linear": (ctypes.c_int, [ctypes.POINTER(MeshView), ctypes.POINTER(FormDesc),
                                          ctypes.c_void_p, ctypes.POINTER(IncidenceView), ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p]),
    "tm_assemble_linear_noalloc": (ctypes.c_int, [ctypes.POINTER(MeshView), ctypes.POINTER(FormDesc),
                                                  ctypes.c_void_p, ctypes.POINTER(IncidenceView), ctypes.c_void_p,
                                                  ctypes.c_void_p]),
    "tm_colpos_build": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int64]
                        + [ctypes.c_void_p] * 6),
    "tm_assemble_p1_rows": (ctypes.c_int, [ctypes.POINTER(MeshView), ctypes.POINTER(FormDesc),
                                           ctypes.POINTER(RowsView), ctypes.c_void_p, ctypes.c_void_p]),
    "tm_rows_dest_build": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(RowsView), ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p]),
    "tm_assemble_p1_adjoint": (ctypes.c_int, [ctypes.POINTER(MeshView), ctypes.POINTER(FormDesc), ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.POINTER(IncidenceView), ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "tm_spmv": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64] + [ctypes.c_void_p] * 6),
    "tm_jacobi": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64] + [ctypes.c_void_p] * 5),
    "tm_pcg_init": (ctypes.c_int, [ctypes.POINTER(PcgView), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "tm_pcg_iterate": (ctypes.c_int, [ctypes.POINTER(PcgView), ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]),
    "tm_assemble_neohookean": (ctypes.c_int, [ctypes.POINTER(MeshView), ctypes.POINTER(FormDesc), ctypes.c_void_p,
                                              ctypes.POINTER(PatternView), ctypes.POINTER(IncidenceView),
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p]),
    "tm_neohookean_energy": (ctypes.c_int, [ctypes.POINTER(MeshView), ctypes.POINTER(FormDesc), ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(IncidenceView),
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "tm_reduce_p1_rows": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(RowsView), ctypes.c_void_p,
                                         ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]),
    "tm_element_matrices_inc": (ctypes.c_int, [ctypes.POINTER(MeshView), ctypes.POINTER(FormDesc), ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p]),
    "tm_pattern_device_count": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int64]
                                + [ctypes.c_void_p] * 5),
    "tm_pattern_device_fill": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int64]
                               + [ctypes.c_void_p] * 10),
    "tm_assemble_p1_incidence": (ctypes.c_int, [ctypes.POINTER(MeshView), ctypes.POINTER(FormDesc),
                                                ctypes.POINTER(RowsView), ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_void_p]),
    "tm_block_rows_pairs": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64] + [ctypes.c_void_p] * 8),
    "tm_assemble_neohookean_rows": (ctypes.c_int, [ctypes.POINTER(MeshView), ctypes.POINTER(FormDesc),
                                                   ctypes.c_void_p, ctypes.POINTER(BlockRowsView)]
                                    + [ctypes.c_void_p] * 6),
    "tm_assemble_facet_linear": (ctypes.c_int, [ctypes.POINTER(MeshView), ctypes.POINTER(FacetView), ctypes.c_void_p,
                                                ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.POINTER(IncidenceView), ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p]),
    "tm_condense": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_int64] + [ctypes.c_void_p] * 12),
    "tm_assemble_j2": (ctypes.c_int, [ctypes.POINTER(MeshView), ctypes.POINTER(FormDesc), ctypes.POINTER(J2Material),
                                      ctypes.c_void_p, ctypes.POINTER(J2State), ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.POINTER(PatternView), ctypes.POINTER(IncidenceView)]
                       + [ctypes.c_void_p] * 5),
    "tm_contact_penalty": (ctypes.c_int, [ctypes.POINTER(MeshView), ctypes.POINTER(FacetView), ctypes.c_void_p,
                                          ctypes.c_int32, ctypes.POINTER(ContactPlane), ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(IncidenceView),
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(PatternView),
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "tm_j2_update_state": (ctypes.c_int, [ctypes.POINTER(MeshView), ctypes.POINTER(FormDesc),
                                          ctypes.POINTER(J2Material), ctypes.c_void_p, ctypes.POINTER(J2State),
                                          ctypes.c_void_p]),
}

_LIB: Optional[ctypes.CDLL] = None


class NativeError(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise NativeError(
            f"{LIB_PATH} is missing: build it with `python -m tensormesh_amd.build` "
            "(the HIP kernels are the only implementation of the built-in forms)")
    handle = ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | getattr(os, "RTLD_LOCAL", 0))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    ver = handle.tm_abi_version()
    if ver != ABI_VERSION:
        raise NativeError(f"libtmhip.so ABI version {ver}, expected {ABI_VERSION}")
    _LIB = handle
    return handle


def check(status: int, what: str) -> None:
    if status != 0:
        msg = lib().tm_last_error().decode(errors="replace")
        raise NativeError(f"{what} failed with status {status}: {msg}")


def ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    if t is None:
        return None
    return t.data_ptr()


def require_device(t: torch.Tensor, what: str) -> None:
    if t.device.type != "cuda":
        raise NativeError(
            f"{what} is on {t.device}; the built-in forms run on the ROCm device "
            "(move the mesh / data with .cuda())")


def stream_of(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


LDS_BYTES = 16384   # LDS staging buffer of the planned segmented reduction (tm_reduce.hip)


def reduce_cap(dtype: torch.dtype, hh: int) -> int:
    """Contributions per reduction block that fit the kernel's LDS buffer."""
    size = 8 if dtype == torch.float64 else 4
    return max(1, min(4096, LDS_BYTES // (size * hh)))


def reduce_plan(seg_ptr_host, cap: int, max_segments: Optional[int] = None):
    """blk_ptr (int32 numpy) of tm_reduce_plan for a host seg_ptr array;
    with max_segments, blocks holding more segments are split."""
    import numpy as np
    seg_ptr_host = np.ascontiguousarray(seg_ptr_host, dtype=np.int32)
    n = seg_ptr_host.shape[0] - 1
    out = np.empty(n + 1, np.int32)
    nb = ctypes.c_int64(0)
    check(lib().tm_reduce_plan(seg_ptr_host.ctypes.data, n, cap, out.ctypes.data, ctypes.byref(nb)),
          "tm_reduce_plan")
    bp = out[: nb.value + 1].copy()
    if max_segments is not None and bp.size > 1 and int(np.diff(bp).max()) > max_segments:
        pieces = [np.arange(a, b, max_segments, dtype=np.int64) for a, b in zip(bp[:-1], bp[1:])]
        bp = np.concatenate(pieces + [np.array([n], np.int64)]).astype(np.int32)
    return bp


def dtype_code(dtype: torch.dtype) -> int:
    if dtype == torch.float64:
        return TM_F64
    if dtype == torch.float32:
        return TM_F32
    raise NativeError(f"dtype {dtype} is not supported (float32 / float64)")
