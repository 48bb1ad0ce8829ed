"""tm_assemble_linear_noalloc (slot values in the library's device array)
against tm_assemble_linear (caller workspace), through the C-ABI: bit-equal
load vectors on two calls in a row, the second on a smaller mesh, in both
precisions, and the first also against the oracle."""
import ctypes
import math

import pytest
import torch

import tensormesh_amd as tm
from oracle import fem
from tensormesh_amd import _native as nat
from tensormesh_amd.mesh import unit_cube_tet, unit_square_tri

DEV = "cuda:0"


def _source(m):
    x, y = m.points[:, 0], m.points[:, 1]
    return 2 * math.pi ** 2 * torch.sin(math.pi * x) * torch.sin(math.pi * y)


def _both(asm, points, f):
    """(tm_assemble_linear, tm_assemble_linear_noalloc) on the same inputs."""
    L = nat.lib()
    dev = points.device
    E, B = asm._conn32(dev).shape
    mesh, form, cval, iv, keep = asm._linear_args(points, f, 1.0)
    st = nat.stream_of(dev)
    slot = torch.empty(E * B, dtype=points.dtype, device=dev)
    ref = torch.full((points.shape[0],), float("nan"), dtype=points.dtype, device=dev)
    out = torch.full((points.shape[0],), float("nan"), dtype=points.dtype, device=dev)
    cv = None if cval is None else ctypes.byref(cval)
    nat.check(L.tm_assemble_linear(ctypes.byref(mesh), ctypes.byref(form), cv, ctypes.byref(iv), slot.data_ptr(),
                                   ref.data_ptr(), st), "tm_assemble_linear")
    nat.check(L.tm_assemble_linear_noalloc(ctypes.byref(mesh), ctypes.byref(form), cv, ctypes.byref(iv),
                                           out.data_ptr(), st), "tm_assemble_linear_noalloc")
    torch.cuda.synchronize()
    return ref, out


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_noalloc_load_equals_workspace_entry_point(dtype):
    cases = [("tetra", unit_cube_tet(12, 0.15, 0)), ("triangle", unit_square_tri(20, 0.2, 0)),
             ("tetra", unit_cube_tet(3, 0.15, 1))]
    for k, (etype, m) in enumerate(cases):
        print(f"noalloc case {k}: {etype} E={m.elements().shape[0]} {dtype}", flush=True)
        md = m.to(DEV)
        pts = md.points.to(dtype).contiguous()
        f = _source(m).to(device=DEV, dtype=dtype).contiguous()
        asm = tm.SourceNodeAssembler.from_mesh(md, quadrature_order=2)
        ref, out = _both(asm, pts, f)
        assert torch.equal(ref, out), f"case {k}: noalloc differs from the workspace entry point"
        if dtype == torch.float64 and k == 0:
            want = fem.assemble_linear(m.points, m.elements(), etype, 2, _source(m))
            assert (out.cpu() - want).abs().max() <= 1e-12 * want.abs().max()
        # constant load c * v through the same array
        mesh, form, _, iv, keep = asm._linear_args(pts, None, 1.0)
        c = ctypes.c_double(0.75)
        a = torch.empty(pts.shape[0], dtype=dtype, device=DEV)
        b = torch.empty(pts.shape[0], dtype=dtype, device=DEV)
        slot = torch.empty(asm._conn32(torch.device(DEV)).numel(), dtype=dtype, device=DEV)
        L = nat.lib()
        st = nat.stream_of(torch.device(DEV))
        nat.check(L.tm_assemble_linear(ctypes.byref(mesh), ctypes.byref(form), ctypes.byref(c), ctypes.byref(iv),
                                       slot.data_ptr(), a.data_ptr(), st), "tm_assemble_linear")
        nat.check(L.tm_assemble_linear_noalloc(ctypes.byref(mesh), ctypes.byref(form), ctypes.byref(c),
                                               ctypes.byref(iv), b.data_ptr(), st), "tm_assemble_linear_noalloc")
        torch.cuda.synchronize()
        assert torch.equal(a, b)
