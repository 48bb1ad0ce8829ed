"""Differentiable P1 assembly with a hand-written adjoint.

``assemble_with_adjoint`` returns the assembled entry values as the output
of a ``torch.autograd.Function`` whose backward calls
``tm_assemble_p1_adjoint``: closed-form dL/dkappa and dL/dx with the element
geometry recomputed (nothing but the inputs is saved).  It stands in for the
reference's autograd through its assembly (tensormesh/assemble/
element_assembler.py:416-441, tensormesh/element/element.py:956-968) for the
P1 built-in forms; other forms use the torch path when gradients are needed.
"""

from __future__ import annotations

import ctypes
from typing import Optional

import torch

from . import _native as nat

ADJOINT_FORMS = ("laplace", "diffusion_elem", "diffusion_node", "mass")


def adjoint_supported(assembler, match) -> bool:
    return assembler.etype in ("triangle", "tetra") and match.form in ADJOINT_FORMS


class _P1Assemble(torch.autograd.Function):
    @staticmethod
    def forward(ctx, points, coeff, assembler, match):
        ctx.assembler, ctx.match = assembler, match
        ctx.save_for_backward(points, coeff)
        return assembler._native_values(match, points, coeff)

    @staticmethod
    def backward(ctx, gvals):
        points, coeff = ctx.saved_tensors
        A, match = ctx.assembler, ctx.match
        want_x = ctx.needs_input_grad[0]
        want_c = ctx.needs_input_grad[1] and coeff is not None
        if not (want_x or want_c):
            return None, None, None, None
        L = nat.lib()
        dev = points.device
        dtype = points.dtype
        pts = points.detach().contiguous()
        cf = None if coeff is None else coeff.detach().to(dtype).contiguous()
        conn = A._conn32(dev)
        E, B = conn.shape
        D = A.dimension
        pat = A._pattern_on(dev)
        tab, tab_dev = A._tables(dev)
        inc_ptr, inc = pat.inc_ptr, pat.inc
        cap = nat.reduce_cap(dtype, 1)
        plan = pat.plan(cap, node=True)
        mesh = A._mesh_view(pts)
        from .assemble import _n_quad
        form = nat.FormDesc(nat.FORM[match.form], _n_quad(A.etype, A.quadrature_order), tab.ctypes.data,
                            tab_dev.data_ptr(), None if cf is None else cf.data_ptr(), 0.0, 0.0)
        iv = nat.IncidenceView(inc_ptr.data_ptr(), inc.data_ptr(), plan.numel() - 1, plan.data_ptr())
        g = gvals.detach().to(dtype).contiguous()
        ws_x = torch.empty(E * B * D, dtype=dtype, device=dev) if want_x else None
        gx = torch.empty(A.n_points, D, dtype=dtype, device=dev) if want_x else None
        node_coef = match.form == "diffusion_node"
        ws_c = torch.empty(E * B, dtype=dtype, device=dev) if (want_c and node_coef) else None
        gc = torch.empty(coeff.shape[0], dtype=dtype, device=dev) if want_c else None
        p = lambda t: None if t is None else t.data_ptr()
        nat.check(L.tm_assemble_p1_adjoint(ctypes.byref(mesh), ctypes.byref(form), pat.slot2nnz.data_ptr(),
                                           g.data_ptr(), ctypes.byref(iv), p(ws_x), p(ws_c), p(gx), p(gc),
                                           nat.stream_of(dev)), "tm_assemble_p1_adjoint")
        if gc is not None:
            gc = gc.to(coeff.dtype)
        return gx, gc, None, None


def assemble_with_adjoint(assembler, match, points: torch.Tensor, coeff: Optional[torch.Tensor]) -> torch.Tensor:
    nat.require_device(points, "points")
    return _P1Assemble.apply(points, coeff, assembler, match)
