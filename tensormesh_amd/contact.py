"""Boundary energies: ``ContactAssembler`` and the built-in ``PenaltyContact``.

Mirrors ContactAssembler (tensormesh/assemble/builtin.py:563-728): a
``FacetAssembler`` whose ``energy(points=None, func=None, point_data=None,
element_data=None, scalar_data=None, batch_size=-1)`` integrates the
per-point density ``element_energy`` (or ``func``) over the selected
boundary facets.  Arguments are matched by name as in the reference:
``element_data`` keys are per-element values (``{key: tensor}`` or
``{key: {etype: tensor}}``), ``scalar_data`` keys are passed as they are,
``point_data`` keys (``x`` included, the positions ``points``) are
interpolated at the facet points with the cell basis.

FxW is that of the configuration the assembler was built on; ``points``
moves only the positions the density sees.  The reference caches FxW at its
first evaluation (transformation.py:1081-1096) and ``update_points`` does not
reset it, so after a first call at the mesh's own points it integrates
against the same FxW.

``PenaltyContact`` is the density of the reference's example
(builtin.py:588-597) -- kappa/2 <-(x_y - obstacle_y)>_+^2 -- generalised to a
plane n . x = offset.  On device tensors its energy, gradient and Hessian run
on the HIP kernel ``tm_contact_penalty``; ``energy(points=x)`` is
differentiable in ``x`` (backward is the HIP gradient).
"""

from __future__ import annotations

import ctypes
import inspect
import math
from typing import Optional, Sequence

import torch
from torch.autograd.function import once_differentiable

from . import _native as nat
from .element import facet_reference_tables, facet_tables
from .facet import FacetAssembler
from .pattern import Pattern, layout_hash
from .sparse import SparseMatrix


class ContactAssembler(FacetAssembler):
    """Energy over boundary facets (builtin.py:563-728); subclasses define
    ``element_energy``."""

    def energy(self, points: Optional[torch.Tensor] = None, func=None, point_data=None, element_data=None,
               scalar_data=None, batch_size: int = -1) -> torch.Tensor:
        t = self._topo
        point_data = {} if point_data is None else dict(point_data)
        element_data = {} if element_data is None else {
            k: (v[t.etype] if isinstance(v, dict) else v) for k, v in element_data.items()}
        scalar_data = {} if scalar_data is None else {k: torch.as_tensor(v) for k, v in scalar_data.items()}
        for key, value in point_data.items():
            assert value.shape[0] == self.n_points, f"the shape of {key} should be [n_point, ...]"
        x = t.points if points is None else points
        point_data["x"] = x
        fn = self.element_energy if func is None else func
        e_dims, q_dims = [], []
        for key in inspect.signature(fn).parameters:
            if key in element_data:
                e_dims.append(0)
                q_dims.append(None)
            elif key in scalar_data:
                e_dims.append(None)
                q_dims.append(None)
            elif key in point_data:
                e_dims.append(0)
                q_dims.append(0)
            else:
                raise ValueError(f"{key} is not supported for contact energy calculation.")
        dev = x.device
        elem = t.elem.to(dev)
        if elem.numel() == 0:
            return torch.zeros((), dtype=x.dtype, device=dev)
        phi, _, fxw = self._geometry(t.points.to(device=dev, dtype=x.dtype))
        dofs = t.conn.to(dev)[elem]
        args = []
        for key in inspect.signature(fn).parameters:
            if key in element_data:
                args.append(element_data[key].to(dev)[elem])
            elif key in scalar_data:
                args.append(scalar_data[key].to(dev))
            else:
                args.append(torch.einsum("sb...,sqb->sq...", point_data[key][dofs], phi))
        density = torch.vmap(torch.vmap(fn, in_dims=tuple(q_dims)), in_dims=tuple(e_dims))(*args)
        return (density * fxw).sum()

    def element_energy(self, **kwargs):
        raise NotImplementedError("element_energy is not implemented")


class _ContactEnergy(torch.autograd.Function):
    """Penalty energy on tm_contact_penalty; when the positions require a
    gradient the same launch assembles dE/dx, which backward scales."""

    @staticmethod
    def forward(ctx, x, model):
        want_grad = ctx.needs_input_grad[0]
        energy, grad, _ = model._native(x, want_energy=True, want_grad=want_grad)
        ctx.save_for_backward(grad if want_grad else None)
        ctx.x_shape, ctx.x_dtype = x.shape, x.dtype
        return energy

    @staticmethod
    @once_differentiable
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return (g * grad).reshape(ctx.x_shape).to(ctx.x_dtype), None


class PenaltyContact(ContactAssembler):
    """kappa/2 <-(n . x - offset)>_+^2 on the boundary facets: the example
    density of ContactAssembler (builtin.py:588-597, the plane y = obstacle_y)
    with a general plane.  ``residual(points)`` / ``tangent(points)`` give
    dE/dx [N*D] and d2E/dx2 as a block SparseMatrix in the BSR(D x D)
    node-pair layout of the cell mesh's pattern (the layout of the element
    assemblers' vector tangents, so the two add)."""

    #: True runs energy through the generic vmapped path (torch ops on the same device).
    reference_path: bool = False

    def __post_init__(self, kappa: float = 1e6, obstacle_y: float = 0.0, normal: Optional[Sequence[float]] = None,
                      offset: Optional[float] = None):
        D = self.dimension
        self.kappa = float(kappa)
        self.obstacle_y = float(obstacle_y)
        n = [0.0] * D
        if normal is None:
            n[1] = 1.0
        else:
            assert len(normal) == D, f"normal must have {D} components"
            n = [float(v) for v in normal]
        s = math.sqrt(sum(v * v for v in n))
        assert s > 0, "normal must be nonzero"
        self.normal = tuple(v / s for v in n)
        self.offset = self.obstacle_y if offset is None else float(offset)

    def element_energy(self, x):
        n = torch.tensor(self.normal, dtype=x.dtype, device=x.device)
        gap = (x * n).sum(-1) - self.offset
        penetration = torch.clamp(-gap, min=0.0)
        return 0.5 * self.kappa * penetration ** 2

    def energy(self, points: Optional[torch.Tensor] = None, func=None, point_data=None, element_data=None,
               scalar_data=None, batch_size: int = -1) -> torch.Tensor:
        x = self._topo.points if points is None else points
        if (func is None and not point_data and not element_data and not scalar_data and x.is_cuda
                and not self.reference_path):
            return _ContactEnergy.apply(x, self)
        return super().energy(points, func, point_data, element_data, scalar_data, batch_size)

    def residual(self, points: Optional[torch.Tensor] = None) -> torch.Tensor:
        return self._native(self._topo.points if points is None else points, want_grad=True)[1]

    def tangent(self, points: Optional[torch.Tensor] = None) -> SparseMatrix:
        return self._native(self._topo.points if points is None else points, want_hess=True)[2]

    def residual_and_tangent(self, points: Optional[torch.Tensor] = None):
        _, r, k = self._native(self._topo.points if points is None else points, want_grad=True, want_hess=True)
        return r, k

    # ---- device data ----
    def _static(self, dev):
        key = f"contact@{dev}"
        if key not in self._dev_cache:
            t = self._topo
            tab = torch.from_numpy(facet_reference_tables(t.etype, t.quadrature_order)).to(dev)
            conn32 = t.conn.to(device=dev, dtype=torch.int32).contiguous()
            elem32 = t.elem.to(device=dev, dtype=torch.int32).contiguous()
            loc32 = t.local.to(device=dev, dtype=torch.int32).contiguous()
            inc_ptr, inc = (a.to(dev) for a in t.incidence)
            nq = int(facet_tables(t.etype, t.quadrature_order)[0].shape[0])
            self._dev_cache[key] = (tab, conn32, elem32, loc32, inc_ptr, inc, nq)
        return self._dev_cache[key]

    def _facet_segments(self, dev):
        """Cell-mesh pattern and, per pattern entry, the facet slots
        f*B*B + b*B + c landing on it (ascending): setup, once per device."""
        key = f"contact_seg@{dev}"
        if key not in self._dev_cache:
            t = self._topo
            conn = t.conn.to(dev)
            pat = Pattern.build_device(conn, self.n_points) if dev.type == "cuda" else None
            if pat is None:
                pat = Pattern.build(conn, self.n_points, device=dev)
            B = conn.shape[1]
            ar = torch.arange(B, device=dev)
            eslot = t.elem.to(dev)[:, None, None] * (B * B) + ar[:, None] * B + ar[None, :]
            entry = pat.slot2nnz.to(dev).long()[eslot.reshape(-1)]
            order = torch.argsort(entry, stable=True)
            counts = torch.bincount(entry, minlength=pat.nnz)
            # only the entries some facet touches: the reduction writes those, the
            # caller zeroes the rest of the Hessian (most of the cell pattern)
            touched = torch.nonzero(counts).reshape(-1)
            seg_ptr = torch.zeros(touched.numel() + 1, dtype=torch.int32, device=dev)
            seg_ptr[1:] = torch.cumsum(counts[touched], 0).to(torch.int32)
            self._dev_cache[key] = (pat, seg_ptr, order.to(torch.int32).contiguous(),
                                    touched.to(torch.int32).contiguous())
        return self._dev_cache[key]

    def _wrap(self, values: torch.Tensor, pat: Pattern, dev) -> SparseMatrix:
        D = self.dimension
        key = f"contact_block@{dev}"
        if key not in self._dev_cache:
            a = torch.arange(D, device=dev).repeat_interleave(D)
            b = torch.arange(D, device=dev).repeat(D)
            r = (pat.row.to(dev)[:, None] * D + a[None]).reshape(-1)
            c = (pat.col.to(dev)[:, None] * D + b[None]).reshape(-1)
            self._dev_cache[key] = (r, c, layout_hash(r.cpu().numpy(), c.cpu().numpy()))
        r, c, h = self._dev_cache[key]
        return SparseMatrix(values.reshape(-1), r, c, (self.n_points * D, self.n_points * D), layout_hash=h)

    def _native(self, x, want_energy=False, want_grad=False, want_hess=False):
        """(energy [], gradient [N*D], Hessian SparseMatrix) on tm_contact_penalty."""
        nat.require_device(x, "points")
        dev, dtype = x.device, x.dtype
        assert tuple(x.shape) == (self.n_points, self.dimension), "points must be [n_points, D]"
        x = x.detach().contiguous()
        ref = self._topo.points.to(device=dev, dtype=dtype).contiguous()
        tab, conn32, elem32, loc32, inc_ptr, inc, nq = self._static(dev)
        F, B, D = int(elem32.numel()), int(conn32.shape[1]), self.dimension
        mesh = nat.MeshView(nat.CELL[self._topo.etype], nat.dtype_code(dtype), conn32.shape[0], ref.shape[0],
                            conn32.data_ptr(), ref.data_ptr())
        fv = nat.FacetView(F, elem32.data_ptr(), loc32.data_ptr())
        n3 = list(self.normal) + [0.0] * (3 - D)
        plane = nat.ContactPlane(self.kappa, (ctypes.c_double * 3)(*n3), self.offset)
        iv = nat.IncidenceView(inc_ptr.data_ptr(), inc.data_ptr(), 0, None)
        new = lambda n, want, dt=dtype: torch.empty(max(n, 1), dtype=dt, device=dev) if want else None
        ews = new(F + nat.J2_ENERGY_PARTS, want_energy, torch.float64)
        energy = torch.empty((), dtype=dtype, device=dev) if want_energy else None
        rslot, grad = new(F * B * D, want_grad), new(self.n_points * D, want_grad)
        pv, pat, tslot, hess = None, None, None, None
        if want_hess:
            pat, seg_ptr, seg_slot, touched = self._facet_segments(dev)
            key = f"contact_plan@{dev}@{dtype}"
            if key not in self._dev_cache:
                bp = nat.reduce_plan(seg_ptr.cpu().numpy(), nat.reduce_cap(dtype, D * D))
                self._dev_cache[key] = torch.from_numpy(bp).to(dev)
            bp = self._dev_cache[key]
            pv = nat.PatternView(touched.numel(), None, None, seg_ptr.data_ptr(), seg_slot.data_ptr(),
                                 bp.numel() - 1, bp.data_ptr(), touched.data_ptr(), touched.data_ptr())
            tslot = new(F * B * B * D * D, True)
            hess = torch.zeros(max(pat.nnz * D * D, 1), dtype=dtype, device=dev)
        nat.check(nat.lib().tm_contact_penalty(ctypes.byref(mesh), ctypes.byref(fv), tab.data_ptr(), nq,
                                               ctypes.byref(plane), x.data_ptr(), nat.ptr(ews), nat.ptr(energy),
                                               ctypes.byref(iv), nat.ptr(rslot), nat.ptr(grad),
                                               ctypes.byref(pv) if pv is not None else None, nat.ptr(tslot),
                                               nat.ptr(hess), nat.stream_of(dev)), "tm_contact_penalty")
        K = self._wrap(hess[:pat.nnz * D * D], pat, dev) if want_hess else None
        return energy, (grad[:self.n_points * D] if want_grad else None), K
