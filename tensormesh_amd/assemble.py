"""ElementAssembler / NodeAssembler with the reference's API.

Mirrors tensormesh/assemble/element_assembler.py:35-879 and
tensormesh/assemble/node_assembler.py:33-741 for meshes of one cell type
(triangle, tetra, triangle6, tetra10, triangle10, tetra20, quad, quad9,
hexahedron, hexahedron27):

* ``from_mesh(mesh, quadrature_order=2, project='reduce', *args, **kwargs)``
  builds the CSR pattern and the slot permutation once (libtmhip host
  builder) and forwards ``*args, **kwargs`` to ``__post_init__``;
* ``from_assembler(obj, *args, **kwargs)`` shares that topology;
* ``__call__(points=None, func=None, point_data=None, element_data=None,
  scalar_data=None, batch_size=-1)`` returns a ``SparseMatrix`` whose
  entries are in the reference's edge order (sorted unique (row, col)), or
  for ``[B, B, H, H]`` integrands the ``from_block_coo`` expansion.

Dispatch: a ``forward`` (or ``func``) recognised as a built-in integrand
(``tensormesh_amd.forms``) runs on the HIP kernels and requires the data on
the ROCm device -- there is no CPU or torch substitute for it.  Any other
integrand runs through the reference's generic vmap path, restated in
``_torch_bilinear`` / ``_torch_linear`` below; so do built-in integrands when
gradients are requested for a form that has no hand-written adjoint yet.

Geometry is recomputed on every call from the points passed in; the
reference's cached-geometry quirk (transformation.py:137-161, stale
``_shape_grad`` after ``update_points``) is deliberately not reproduced.
"""

from __future__ import annotations

import ctypes
import inspect
import math
from typing import Callable, Dict, Mapping, Optional, Union

import numpy as np
import torch
import torch.nn as nn

from . import _native as nat
from .element import ELEMENT_SHAPE, reference_tables, torch_tables
from .forms import FormMatch, MatchCache, match_bilinear, match_linear
from .pattern import Pattern, layout_hash
from .sparse import SparseMatrix


# ------------------------------------------------------------ geometry ---
def shape_grad_jxw(points: torch.Tensor, conn: torch.Tensor, etype: str, qorder: int,
                   q_start: int = 0, q_batch: int = -1):
    """Torch evaluation of the reference geometry (element.py:956-968,
    transformation.py:1265-1272) for the generic form path."""
    w, phi, dphi = torch_tables(etype, qorder, points.dtype, points.device)
    stop = None if q_batch in (-1, None) else q_start + q_batch
    w, phi, dphi = w[q_start:stop], phi[q_start:stop], dphi[q_start:stop]
    coords = points[conn]
    J = torch.einsum("ebj,qib->eqij", coords, dphi)
    inv = torch.inverse(J)
    grad = torch.einsum("qib,eqji->eqbj", dphi, inv)
    jxw = torch.einsum("q,eq->eq", w, torch.linalg.det(J).abs())
    return phi, grad, jxw


def _n_quad(etype: str, qorder: int) -> int:
    from .element import quadrature
    return int(quadrature(etype, qorder)[0].shape[0])


class _Broadcast:
    def __init__(self, element, quadrature, u, v):
        self.element = 0 if element else None
        self.quadrature = 0 if quadrature else None
        self.u = 0 if u else None
        self.v = 0 if v else None


# --------------------------------------------------------- base module ---
class _AssemblerBase(nn.Module):
    def _init_topology(self, elements: Dict[str, torch.Tensor], points: torch.Tensor, quadrature_order: int):
        if len(elements) != 1:
            raise NotImplementedError("tensormesh_amd assembles one element type per mesh")
        etype = next(iter(elements))
        if etype not in ELEMENT_SHAPE:
            raise NotImplementedError(f"element type {etype} is not supported ({', '.join(ELEMENT_SHAPE)})")
        self.element_types = [etype]
        self.elements = {etype: elements[etype].long()}
        self.dimension = ELEMENT_SHAPE[etype][0]
        self.quadrature_order = quadrature_order
        self.n_points = int(points.shape[0])
        self._points = points
        self._dev_cache: Dict[str, object] = {}

    @property
    def etype(self) -> str:
        return self.element_types[0]

    @property
    def device(self) -> torch.device:
        return self._points.device

    @property
    def dtype(self) -> torch.dtype:
        return self._points.dtype

    def type(self, dtype: torch.dtype):
        if dtype not in (torch.float32, torch.float64):
            raise Exception(f"the dtype {dtype} is not supported")
        self._points = self._points.to(dtype)
        return self

    def _cached(self, key: str, device: torch.device, make):
        k = f"{key}@{device}"
        if k not in self._dev_cache:
            self._dev_cache[k] = make()
        return self._dev_cache[k]

    def _conn32(self, device) -> torch.Tensor:
        return self._cached("conn32", device,
                            lambda: self.elements[self.etype].to(device=device, dtype=torch.int32).contiguous())

    def _tables(self, device):
        tab = reference_tables(self.etype, self.quadrature_order)
        dev = self._cached(f"tab{self.quadrature_order}", device,
                           lambda: torch.from_numpy(tab).to(device))
        return tab, dev

    def _mesh_view(self, points: torch.Tensor) -> nat.MeshView:
        conn = self._conn32(points.device)
        return nat.MeshView(nat.CELL[self.etype], nat.dtype_code(points.dtype), conn.shape[0], points.shape[0],
                            conn.data_ptr(), points.data_ptr())


# ----------------------------------------------------- ElementAssembler ---
class ElementAssembler(_AssemblerBase):
    """Bilinear-form assembler; see the module docstring."""

    #: P1 scalar forms: "rows" = fused row-centric kernel (tm_assemble_p1_rows),
    #: "two_stage" = element kernel + segmented reduction (tm_assemble_bilinear),
    #: "two_stage_rows" = element kernel + row-centric reduction (tm_reduce_p1_rows),
    #: "incidence" = element kernel writing in incidence order + contiguous row
    #: reduction (tm_assemble_p1_incidence), "auto" = the fastest of the four on
    #: this device and pattern, timed once and cached (tensormesh_amd.tune).
    p1_path: str = "rows"
    #: True: symmetric built-in forms on the element-matrix, Neo-Hookean and
    #: J2 kernels compute and store only the element blocks (i, j) with i <= j
    #: and the reduction reads the others transposed (TM_FLAG_UPPER with the
    #: pattern's symmetric segment list); the products are formed so that the
    #: assembled values are bit for bit those of the full blocks.  False
    #: stores every block, as the reference does.  None (default) picks per
    #: kernel what measured faster on MI355X: symmetric for the D x D block
    #: forms (linear elasticity, the Neo-Hookean and J2 tangents), whose
    #: element kernels are bound by their block stores; full blocks for scalar
    #: forms, where the transposed gathers cost more than the halved element
    #: work saves.
    symmetric_slots: Optional[bool] = None
    #: Float32 meshes on the P1 scalar kernels evaluate element geometry in
    #: float32 and redo elements of shape quality q = |det J| / prod |edges|
    #: below 0.1 in float64 (the float32 rounding error of an element matrix
    #: grows like eps / q).  False keeps float32 for every element
    #: (TM_FLAG_F32_NO_REDO; a diagnostic for tests of that criterion).
    sliver_redo: bool = True
    #: Build the pattern on the GPU (tm_pattern_device_*) when the mesh is on the device.
    device_pattern: bool = True
    #: True runs every integrand through the restated reference vmap/autograd
    #: path (torch ops on the same device) -- the "plain" side of a
    #: measurement, never chosen implicitly.
    reference_path: bool = False
    #: Elements per wavefront of the wave kernels (element matrices of P2 and
    #: block forms, Neo-Hookean, J2): None = each kernel's own choice, an int =
    #: that count (the kernel rejects one the cell, rule or LDS budget does not
    #: admit), "auto" = the fastest admitted count, timed once per device,
    #: kernel, cell, form, dtype and rule and cached (tensormesh_amd.tune).
    #: The results do not depend on it.
    epw: Union[int, str, None] = None

    def _launch_epw(self, kernel: str, dev, dtype, tag: str, run: Callable[[int], object], options) -> int:
        """tm_form_desc.epw for a launch: 0 (the kernel's own choice), the
        fixed ``epw``, or the tuned fastest of ``options`` (``run(e)`` makes the
        call with e elements per wavefront)."""
        if self.epw is None:
            return 0
        if self.epw != "auto":
            return int(self.epw)
        from .tune import device_tag, tuner
        key = f"{device_tag(dev)}|epw|{kernel}|{self.etype}|{tag}|{dtype}|q{self.quadrature_order}"

        def fixed(e):
            saved = self.epw
            self.epw = e
            try:
                return run(e)
            finally:
                self.epw = saved
        return int(tuner().choose(key, {str(e): (lambda e=e: fixed(e)) for e in options}, device=dev))

    def __init__(self, elements: Dict[str, torch.Tensor], points: torch.Tensor, pattern: Pattern,
                 quadrature_order: int = 2, *args, **kwargs):
        super().__init__()
        self._init_topology(elements, points, quadrature_order)
        self._pattern = pattern
        self.edges = torch.stack([pattern.row, pattern.col], 0)
        self.__post_init__(*args, **kwargs)

    def __post_init__(self, *args, **kwargs):
        pass

    def forward(self, *args, **kwargs):
        raise NotImplementedError("forward is not implemented")

    # ---- constructors (element_assembler.py:741-878) ----
    @classmethod
    def from_mesh(cls, mesh, quadrature_order: int = 2, project: str = "reduce", *args, **kwargs):
        assert project in ["reduce", "sparse"]
        elements = mesh.elements()
        if isinstance(elements, torch.Tensor):
            elements = {mesh.default_element_type: elements}
        if len(elements) != 1:
            raise NotImplementedError("tensormesh_amd assembles one element type per mesh")
        conn = next(iter(elements.values()))
        pattern = None
        if mesh.points.device.type == "cuda" and cls.device_pattern:
            pattern = Pattern.build_device(conn.to(mesh.points.device), mesh.points.shape[0])
        if pattern is None:
            pattern = Pattern.build(conn, mesh.points.shape[0], device=mesh.points.device)
        return cls(elements, mesh.points, pattern, quadrature_order, *args, **kwargs)

    @classmethod
    def from_assembler(cls, obj, *args, **kwargs):
        assert isinstance(obj, ElementAssembler), f"obj must be an instance of ElementAssembler, but got {type(obj)}"
        return cls(obj.elements, obj._points, obj._pattern, obj.quadrature_order, *args, **kwargs)

    @property
    def pattern(self) -> Pattern:
        return self._pattern

    def _pattern_on(self, device) -> Pattern:
        if self._pattern.row_ptr.device == torch.device(device) or (
                torch.device(device).type == "cuda" and self._pattern.row_ptr.device.type == "cuda"):
            return self._pattern
        return self._cached("pattern", torch.device(device), lambda: self._pattern.to(device))

    # ---- __call__ (element_assembler.py:231-445) ----
    def __call__(self, points: Optional[torch.Tensor] = None, func: Optional[Callable] = None,
                 point_data: Optional[Mapping[str, torch.Tensor]] = None,
                 element_data: Optional[Mapping] = None,
                 scalar_data: Optional[Mapping] = None, batch_size: int = -1) -> SparseMatrix:
        assert isinstance(point_data, dict) or point_data is None, \
            f"point_data should be a dict, but got {type(point_data)}. Please pass  in extra parameter using key-value pairs"
        point_data = {} if point_data is None else dict(point_data)
        E = self.elements[self.etype].shape[0]
        if element_data is None:
            element_data = {}
        else:
            element_data = dict(element_data)
            for key, value in element_data.items():
                if isinstance(value, dict):
                    value = value[self.etype]
                    element_data[key] = value
                assert value.shape[0] == E, f"the shape of {key} should be [{E}, ...], but got {value.shape[0]}"
        scalar_data = {} if scalar_data is None else {k: torch.as_tensor(v) for k, v in scalar_data.items()}
        if points is None:
            points = self._points
        else:
            assert points.shape[1] == self.dimension, \
                f"the dimension of points should be {self.dimension}, but got {points.shape[1]}"
        if points.dtype not in (torch.float32, torch.float64):
            raise Exception(f"the dtype {points.dtype} is not supported")
        for key, value in point_data.items():
            assert value.shape[0] == points.shape[0], f"the shape of {key} should be [n_point, ...], but got {value.shape}"

        fn = self.forward if func is None else func
        # the device decides the path: tensors on the ROCm device take the HIP
        # kernels when the form is built in; CPU tensors run the reference's
        # generic torch path, as the reference itself does on the CPU
        on_device = points.device.type == "cuda"
        match = None if (self.reference_path or not on_device) else \
            self._match(fn, func is None, element_data, point_data, scalar_data)
        if match is not None:
            coeff = None
            if match.coeff_key is not None:
                coeff = (element_data if match.form == "diffusion_elem" else point_data)[match.coeff_key]
            needs_grad = torch.is_grad_enabled() and (points.requires_grad or (coeff is not None and coeff.requires_grad))
            if not needs_grad:
                return self._native_bilinear(match, points, coeff)
            from .autograd import adjoint_supported, assemble_with_adjoint
            if adjoint_supported(self, match):
                return self._wrap_output(assemble_with_adjoint(self, match, points, coeff), points.device)
        point_data["x"] = points
        return self._torch_bilinear(fn, points, point_data, element_data, scalar_data, batch_size)

    def _match(self, fn, is_forward: bool, element_data, point_data, scalar_data) -> Optional[FormMatch]:
        """Form recognition, cached per (integrand, its free state, data
        layout): the probe runs again only when something the integrand can
        read changes (``tensormesh_amd.forms.MatchCache``)."""
        def sig(d):
            return tuple(sorted((k, tuple(v.shape[1:]) if torch.is_tensor(v) else None) for k, v in d.items()))

        layout = (is_forward, sig(element_data), sig(point_data), tuple(sorted(scalar_data)))
        cache = self.__dict__.setdefault("_match_cache", MatchCache())
        return cache.get(fn, layout, lambda: match_bilinear(fn, self.dimension, element_data, point_data,
                                                            scalar_data, owner=self if is_forward else None))

    # ---- native path ----
    def _native_values(self, match: FormMatch, points: torch.Tensor, coeff: Optional[torch.Tensor]) -> torch.Tensor:
        nat.require_device(points, "points")
        L = nat.lib()
        dev = points.device
        points = points.detach().contiguous()
        dtype = points.dtype
        pat = self._pattern_on(dev)
        conn = self._conn32(dev)
        E, B = conn.shape
        H = self.dimension if match.form == "linear_elasticity" else 1
        if coeff is not None:
            nat.require_device(coeff, match.coeff_key)
            coeff = coeff.detach().to(dtype).contiguous()
        if E == 0 or pat.nnz == 0:
            values = torch.zeros(pat.nnz * H * H, dtype=dtype, device=dev)
            return values if H == 1 else values.reshape(pat.nnz, H, H)
        tab, tab_dev = self._tables(dev)
        mesh = self._mesh_view(points)
        form = nat.FormDesc(nat.FORM[match.form], _n_quad(self.etype, self.quadrature_order),
                            tab.ctypes.data, tab_dev.data_ptr(),
                            None if coeff is None else coeff.data_ptr(), match.young, match.poisson)
        if not self.sliver_redo:
            form.flags |= nat.TM_FLAG_F32_NO_REDO
        path = self.p1_path
        p1_scalar = H == 1 and self.etype in ("triangle", "tetra")
        if p1_scalar and path == "auto":
            from .tune import device_tag, tuner
            key = f"{device_tag(dev)}|p1_path|{self.etype}|{match.form}|{dtype}|{pat.hash}"
            cands = {p: (lambda p=p: self._p1_launch(p, mesh, form, pat, conn, dtype, dev))
                     for p in ("rows", "two_stage", "two_stage_rows", "incidence")}
            path = tuner().choose(key, cands, device=dev)
        if p1_scalar and path in ("rows", "two_stage_rows", "incidence"):
            values = self._p1_launch(path, mesh, form, pat, conn, dtype, dev)
            if values is not None:
                return values
        # the element-matrix (wave) kernel serves everything but P1 scalar forms
        sym = not p1_scalar and self._symmetric_blocks(H > 1)
        if sym:
            form.flags |= nat.TM_FLAG_UPPER
        if not p1_scalar:
            form.epw = self._launch_epw("wave", dev, dtype, f"{match.form}|{int(sym)}",
                                        lambda e: self._native_values(match, points, coeff), (1, 2))
        pairs = B * (B + 1) // 2 if sym else B * B      # packed upper blocks, or all
        slot = torch.empty(E * pairs * H * H, dtype=dtype, device=dev)
        values = torch.empty(pat.nnz * H * H, dtype=dtype, device=dev)
        pv = pat.view(dtype, H * H, symmetric=sym)
        nat.check(L.tm_assemble_bilinear(ctypes.byref(mesh), ctypes.byref(form), ctypes.byref(pv),
                                         slot.data_ptr(), values.data_ptr(), nat.stream_of(dev)),
                  "tm_assemble_bilinear")
        return values if H == 1 else values.reshape(pat.nnz, H, H)

    def _p1_launch(self, path: str, mesh, form, pat, conn, dtype, dev) -> Optional[torch.Tensor]:
        """CSR values of a P1 scalar form on one of the four paths; None when the
        row paths do not apply to this pattern (repeated vertices, long rows)."""
        L = nat.lib()
        E, B = conn.shape
        st = nat.stream_of(dev)
        values = torch.empty(pat.nnz, dtype=dtype, device=dev)
        if path == "two_stage":
            slot = torch.empty(E * B * B, dtype=dtype, device=dev)
            pv = pat.view(dtype, 1)
            nat.check(L.tm_assemble_bilinear(ctypes.byref(mesh), ctypes.byref(form), ctypes.byref(pv),
                                             slot.data_ptr(), values.data_ptr(), st), "tm_assemble_bilinear")
            return values
        rv = pat.rows_view(conn)
        if rv is None:
            return None
        if path == "rows":
            nat.check(L.tm_assemble_p1_rows(ctypes.byref(mesh), ctypes.byref(form), ctypes.byref(rv),
                                            values.data_ptr(), st), "tm_assemble_p1_rows")
        elif path == "incidence":
            slot = torch.empty(E * B * B, dtype=dtype, device=dev)
            nat.check(L.tm_assemble_p1_incidence(ctypes.byref(mesh), ctypes.byref(form), ctypes.byref(rv),
                                                 pat.inc_inv().data_ptr(), slot.data_ptr(), values.data_ptr(), st),
                      "tm_assemble_p1_incidence")
        else:
            slot = torch.empty(E * B * B, dtype=dtype, device=dev)
            nat.check(L.tm_element_matrices(ctypes.byref(mesh), ctypes.byref(form), slot.data_ptr(), st),
                      "tm_element_matrices")
            nat.check(L.tm_reduce_p1_rows(nat.dtype_code(dtype), nat.CELL[self.etype], ctypes.byref(rv),
                                          slot.data_ptr(), 0, values.data_ptr(), st), "tm_reduce_p1_rows")
        return values

    def _symmetric_blocks(self, default: bool) -> bool:
        """Whether a kernel stores symmetric element blocks (symmetric_slots)."""
        return default if self.symmetric_slots is None else bool(self.symmetric_slots)

    def _native_hyperelastic(self, u: torch.Tensor, points: Optional[torch.Tensor], young: float, poisson: float,
                             want_residual: bool = True, want_tangent: bool = True, path: Optional[str] = None):
        """Neo-Hookean residual [N*D] and tangent (block SparseMatrix) at
        displacement u [N, D] on the HIP kernels (tm_assemble_neohookean)."""
        points = self._points if points is None else points
        nat.require_device(points, "points")
        nat.require_device(u, "displacement")
        L = nat.lib()
        dev = points.device
        dtype = points.dtype
        points = points.detach().contiguous()
        u = u.detach().to(dtype).contiguous()
        assert u.shape == points.shape, f"displacement must be [{points.shape[0]}, {points.shape[1]}], got {tuple(u.shape)}"
        pat = self._pattern_on(dev)
        conn = self._conn32(dev)
        E, B = conn.shape
        D = self.dimension
        tab, tab_dev = self._tables(dev)
        mesh = self._mesh_view(points)
        form = nat.FormDesc(0, _n_quad(self.etype, self.quadrature_order), tab.ctypes.data, tab_dev.data_ptr(),
                            None, float(young), float(poisson),
                            nat.TM_FLAG_MFMA if (getattr(self, "use_mfma", False) and dtype == torch.float32) else 0)
        # the block-row reduction is built for simplex cells; other cells use the slot path
        simplex = self.etype in ("triangle", "tetra", "triangle6", "tetra10")
        path = path or getattr(self, "nh_path", "slots")
        if path == "auto" and simplex and want_tangent:
            from .tune import device_tag, tuner
            key = (f"{device_tag(dev)}|nh_path|{self.etype}|{dtype}|{int(form.flags)}|{want_residual}|"
                   f"{pat.hash}")
            cands = {p: (lambda p=p: self._native_hyperelastic(u, points, young, poisson, want_residual, True, p))
                     for p in ("slots", "rows")}
            path = tuner().choose(key, cands, device=dev)
        use_rows = path == "rows" and simplex
        brv = self._block_rows(pat, conn, dev, D) if use_rows else None
        if brv is not None:
            tslot = torch.empty(E * B * B * D * D, dtype=dtype, device=dev) if want_tangent else None
            rslot = torch.empty(E * B * D, dtype=dtype, device=dev) if want_residual else None
            tan = torch.empty(pat.nnz * D * D, dtype=dtype, device=dev) if want_tangent else None
            res = torch.empty(self.n_points * D, dtype=dtype, device=dev) if want_residual else None
            p = lambda t: None if t is None else t.data_ptr()
            nat.check(L.tm_assemble_neohookean_rows(ctypes.byref(mesh), ctypes.byref(form), u.data_ptr(),
                                                    ctypes.byref(brv), pat.inc_inv().data_ptr(), p(tslot), p(tan),
                                                    p(rslot), p(res), nat.stream_of(dev)),
                      "tm_assemble_neohookean_rows")
            K = self._wrap_output(tan.reshape(pat.nnz, D, D), dev) if want_tangent else None
            return res, K
        # the MFMA tile forms every block at once (and not bitwise symmetric): full rows
        sym = want_tangent and not (form.flags & nat.TM_FLAG_MFMA) and self._symmetric_blocks(True)
        if sym:
            form.flags |= nat.TM_FLAG_UPPER
        form.epw = self._launch_epw("neohookean", dev, dtype, f"{int(form.flags)}|{int(want_residual)}|"
                                    f"{int(want_tangent)}",
                                    lambda e: self._native_hyperelastic(u, points, young, poisson, want_residual,
                                                                        want_tangent, "slots"), (1, 2, 4))
        pv = pat.view(dtype, D * D, symmetric=sym)
        node_plan = pat.plan(nat.reduce_cap(dtype, D), node=True)
        iv = nat.IncidenceView(pat.inc_ptr.data_ptr(), pat.inc.data_ptr(), node_plan.numel() - 1, node_plan.data_ptr())
        pairs = B * (B + 1) // 2 if sym else B * B      # packed upper blocks, or all
        tslot = torch.empty(E * pairs * D * D, dtype=dtype, device=dev) if want_tangent else None
        tan = torch.empty(pat.nnz * D * D, dtype=dtype, device=dev) if want_tangent else None
        rslot = torch.empty(E * B * D, dtype=dtype, device=dev) if want_residual else None
        res = torch.empty(self.n_points * D, dtype=dtype, device=dev) if want_residual else None
        p = lambda t: None if t is None else t.data_ptr()
        nat.check(L.tm_assemble_neohookean(ctypes.byref(mesh), ctypes.byref(form), u.data_ptr(), ctypes.byref(pv),
                                           ctypes.byref(iv), p(tslot), p(tan), p(rslot), p(res),
                                           nat.stream_of(dev)), "tm_assemble_neohookean")
        K = self._wrap_output(tan.reshape(pat.nnz, D, D), dev) if want_tangent else None
        return res, K

    def _block_rows(self, pat, conn: torch.Tensor, dev, D: int):
        """Block-row plan and per-contribution pairs for the block-row
        reduction (once per pattern), or None when it cannot serve this mesh
        (a repeated vertex, or a block whose pairs overflow 16 bits)."""
        def make():
            from .pattern import repeated_vertex
            if pat.inc is None or repeated_vertex(conn):
                return None
            E, B = conn.shape
            rp = pat.row_ptr.cpu().numpy()
            ip = pat.inc_ptr.cpu().numpy().astype(np.int64)
            # (entry, component) accumulators per workgroup: 1024 (one pass of the
            # block-row reduction); residual threads: 256 // D rows
            bp = nat.reduce_plan(rp, 1024 // (D * D), max_segments=256 // D)
            t_first = ip[bp[:-1]]
            if int((ip[bp[1:]] - t_first).max(initial=0)) * B > 0xffff:
                return None
            info = np.zeros((bp.size, 4), np.int32)
            info[:, 0] = bp
            info[:, 1] = ip[bp]
            info[:, 2] = rp[bp]
            info[:, 3] = pat.seg_ptr.cpu().numpy()[rp[bp]]
            row_t0 = np.repeat(t_first, np.diff(bp)).astype(np.int32)
            info_d = torch.from_numpy(info).to(dev)
            row_t0_d = torch.from_numpy(row_t0).to(dev)
            pairs = torch.empty(E * B * B, dtype=torch.int16, device=dev)
            flag = torch.zeros(1, dtype=torch.int32, device=dev)
            nat.check(nat.lib().tm_block_rows_pairs(B, self.n_points, pat.row_ptr.data_ptr(), pat.seg_ptr.data_ptr(),
                                                    pat.seg_slot.data_ptr(), pat.inc_inv().data_ptr(),
                                                    row_t0_d.data_ptr(), pairs.data_ptr(), flag.data_ptr(),
                                                    nat.stream_of(dev)), "tm_block_rows_pairs")
            if int(flag.item()) != 0:
                return None
            view = nat.BlockRowsView(pat.nnz, pat.row_ptr.data_ptr(), pat.seg_ptr.data_ptr(), pairs.data_ptr(),
                                     pat.inc_ptr.data_ptr(), pat.inc.data_ptr(), bp.size - 1, info_d.data_ptr())
            view._keep = (pairs, info_d)
            return view
        return self._cached("nh_rows", dev, make)

    def _native_bilinear(self, match, points, coeff) -> SparseMatrix:
        return self._wrap_output(self._native_values(match, points, coeff), points.device)

    def _wrap_output(self, values: torch.Tensor, device) -> SparseMatrix:
        pat = self._pattern_on(device)
        if values.dim() == 1:
            # digest of the scipy int32 edges (pattern.edge_index_dtype), on first use
            return SparseMatrix(values, pat.row, pat.col, (self.n_points, self.n_points),
                                layout_hash=lambda: pat.hash, row_ptr=pat.row_ptr, index_dtype=pat.edge_dtype)
        H = values.shape[1]

        def make_block():
            # from_block_coo index expansion (matrix.py:367-378), once per pattern
            a = torch.arange(H, device=pat.row.device).repeat_interleave(H)
            b = torch.arange(H, device=pat.row.device).repeat(H)
            r = (pat.row[:, None] * H + a[None]).reshape(-1)
            c = (pat.col[:, None] * H + b[None]).reshape(-1)
            return r, c, layout_hash(r.cpu().numpy(), c.cpu().numpy())

        r, c, h = self._cached(f"block{H}", torch.device(device), make_block)
        return SparseMatrix(values.reshape(-1), r, c, (self.n_points * H, self.n_points * H), layout_hash=h)

    # ---- generic torch path (element_assembler.py:320-445) ----
    def _torch_bilinear(self, fn, points, point_data, element_data, scalar_data, batch_size) -> SparseMatrix:
        etype = self.etype
        conn = self.elements[etype].to(points.device)
        params = list(inspect.signature(fn).parameters)
        rules = [
            (lambda x: x == "u", _Broadcast(False, True, True, False)),
            (lambda x: x == "v", _Broadcast(False, True, False, True)),
            (lambda x: x == "gradu", _Broadcast(True, True, True, False)),
            (lambda x: x == "gradv", _Broadcast(True, True, False, True)),
            (lambda x: x in element_data, _Broadcast(True, False, False, False)),
            (lambda x: x in scalar_data, _Broadcast(True, True, True, True)),
            (lambda x: x in point_data, _Broadcast(True, True, False, False)),
            (lambda x: x in {"grad" + k for k in point_data}, _Broadcast(True, True, False, False)),
        ]
        dims = []
        for key in params:
            for cond, bc in rules:
                if cond(key):
                    dims.append(bc)
                    break
            else:
                raise ValueError(f"{key} is not supported, please use `u`, `v`, `gradu`, `gradv` or more keys "
                                 f"provided by point_data, element_data or scalar_data")
        e_dims = tuple(d.element for d in dims)
        q_dims = tuple(d.quadrature for d in dims)
        u_dims = tuple(d.u for d in dims)
        v_dims = tuple(d.v for d in dims)
        inner = torch.vmap(torch.vmap(torch.vmap(fn, in_dims=v_dims), in_dims=u_dims), in_dims=q_dims)
        element_parallel = not all(d is None for d in e_dims)
        pfn = torch.vmap(inner, in_dims=e_dims) if element_parallel else inner
        nq = _n_quad(etype, self.quadrature_order)
        if batch_size in (-1, None):
            n_batch, bs = 1, nq
        else:
            bs, n_batch = batch_size, math.ceil(nq / batch_size)
        ele_pd = {k: v[conn] for k, v in point_data.items()}
        total = None
        for i in range(n_batch):
            phi, grad, jxw = shape_grad_jxw(points, conn, etype, self.quadrature_order, i * bs, bs)
            args = []
            for key in params:
                if key in ("u", "v"):
                    args.append(phi)
                elif key in ("gradu", "gradv"):
                    args.append(grad)
                elif key in ele_pd:
                    args.append(torch.einsum("eb...,qb->eq...", ele_pd[key], phi))
                elif key.startswith("grad") and key[4:] in ele_pd:
                    args.append(torch.einsum("eb...,eqbd->eq...d", ele_pd[key[4:]], grad))
                elif key in element_data:
                    args.append(element_data[key])
                elif key in scalar_data:
                    args.append(scalar_data[key])
            out = pfn(*args)
            if element_parallel:
                out = torch.einsum("eqij...,eq->eij...", out, jxw)
            else:
                out = torch.einsum("qij...,eq->eij...", out, jxw)
            total = out if total is None else total + out
        pat = self._pattern_on(points.device)
        S = total.shape[0] * total.shape[1] * total.shape[2]
        flat = total.reshape(S, *total.shape[3:])
        values = torch.zeros(pat.nnz, *total.shape[3:], dtype=flat.dtype, device=flat.device)
        values = values.index_add(0, pat.slot2nnz.long(), flat)
        if values.dim() == 1:
            return self._wrap_output(values, points.device)
        if values.dim() == 3:
            return self._wrap_output(values, points.device)
        raise Exception(f"the shape of integral is supposed to be  1D or 3D, but got {values.shape}")

    # ---- energy (element_assembler.py:447-659), torch path ----
    def energy(self, points=None, func=None, point_data=None, element_data=None, scalar_data=None,
               batch_size: int = -1):
        point_data = {} if point_data is None else dict(point_data)
        # element_data as {key: tensor} or {key: {etype: tensor}} (element_assembler.py:533-546)
        element_data = {} if element_data is None else {
            k: (v[self.etype] if isinstance(v, dict) else v) for k, v in element_data.items()}
        for k, v in element_data.items():
            assert v.shape[0] == self.elements[self.etype].shape[0], f"element_data[{k!r}] has {v.shape[0]} rows"
        scalar_data = {} if scalar_data is None else {k: torch.as_tensor(v) for k, v in scalar_data.items()}
        points = self._points if points is None else points
        point_data["x"] = points
        fn = self.element_energy if func is None else func
        params = list(inspect.signature(fn).parameters)
        conn = self.elements[self.etype].to(points.device)
        e_dims, q_dims = [], []
        nq = _n_quad(self.etype, self.quadrature_order)
        for key in params:
            if key in element_data:
                d = element_data[key]
                e_dims.append(0)
                q_dims.append(0 if (d.dim() >= 2 and d.shape[1] == nq) else None)
            elif key in scalar_data:
                e_dims.append(None)
                q_dims.append(None)
            elif key in point_data or (key.startswith("grad") and key[4:] in point_data):
                e_dims.append(0)
                q_dims.append(0)
            else:
                raise ValueError(f"{key} is not supported for energy calculation.")
        pfn = torch.vmap(torch.vmap(fn, in_dims=tuple(q_dims)), in_dims=tuple(e_dims))
        phi, grad, jxw = shape_grad_jxw(points, conn, self.etype, self.quadrature_order)
        ele_pd = {k: v[conn] for k, v in point_data.items()}
        args = []
        for key in params:
            if key in ele_pd:
                args.append(torch.einsum("eb...,qb->eq...", ele_pd[key], phi))
            elif key.startswith("grad") and key[4:] in ele_pd:
                args.append(torch.einsum("eb...,eqbd->eq...d", ele_pd[key[4:]], grad))
            elif key in element_data:
                args.append(element_data[key])
            else:
                args.append(scalar_data[key])
        return (pfn(*args) * jxw).sum()

    def element_energy(self, **kwargs):
        raise NotImplementedError("element_energy is not implemented")

    def __repr__(self):
        return (f"{self.__class__.__name__}(element_types={self.element_types}, "
                f"n_element={self.elements[self.etype].shape[0]}, n_point={self.n_points}, "
                f"n_dim={self.dimension}, nnz={self._pattern.nnz})")


# -------------------------------------------------------- NodeAssembler ---
class NodeAssembler(_AssemblerBase):
    """Linear-form assembler (node_assembler.py:33-741)."""

    def __init__(self, elements: Dict[str, torch.Tensor], points: torch.Tensor, incidence, quadrature_order: int = 2,
                 *args, **kwargs):
        super().__init__()
        self._init_topology(elements, points, quadrature_order)
        self._incidence = incidence  # (inc_ptr, inc) int32 on the mesh device
        self._compile = False
        self.__post_init__(*args, **kwargs)

    def __post_init__(self, *args, **kwargs):
        pass

    def forward(self, *args):
        raise NotImplementedError("forward is not implemented")

    @classmethod
    def from_mesh(cls, mesh, quadrature_order: int = 2, project: str = "reduce", *args, **kwargs):
        elements = mesh.elements()
        if isinstance(elements, torch.Tensor):
            elements = {mesh.default_element_type: elements}
        return cls.from_elements(mesh.points, elements, quadrature_order, mesh.device, mesh.dtype, project,
                                 *args, **kwargs)

    @classmethod
    def from_elements(cls, points, elements, quadrature_order: int = 2, device="cpu", dtype=torch.float32,
                      project: str = "reduce", *args, **kwargs):
        if project not in ("reduce", "sparse"):
            raise ValueError(f"project should be either 'reduce' or 'sparse', but got {project}")
        if len(elements) != 1:
            raise NotImplementedError("tensormesh_amd assembles one element type per mesh")
        conn = next(iter(elements.values()))
        L = nat.lib()
        c = np.ascontiguousarray(conn.detach().cpu().numpy().astype(np.int32))
        E, B = c.shape
        n = int(points.shape[0])
        inc_ptr = np.empty(n + 1, np.int32)
        inc = np.empty(E * B, np.int32)
        nat.check(L.tm_incidence_build(c.ctypes.data, E, B, n, inc_ptr.ctypes.data, inc.ctypes.data),
                  "tm_incidence_build")
        dev = torch.device(device)
        incidence = (torch.from_numpy(inc_ptr).to(dev), torch.from_numpy(inc).to(dev))
        return cls(elements, points.to(device=dev, dtype=dtype), incidence, quadrature_order, *args, **kwargs)

    @classmethod
    def from_assembler(cls, obj, *args, **kwargs):
        assert isinstance(obj, NodeAssembler), f"the object {obj} should inheritate from NodeAssembler"
        return cls(obj.elements, obj._points, obj._incidence, obj.quadrature_order, *args, **kwargs)

    def compile(self, mode: str = "disable", **kwargs) -> "NodeAssembler":
        """Accepted for API compatibility (node_assembler.py:416-500): the
        built-in loads are compiled HIP kernels already."""
        self._compile = True
        return self

    def flat_mode(self) -> "NodeAssembler":
        return self.compile()

    def reset_compile(self) -> "NodeAssembler":
        self._compile = False
        return self

    @property
    def is_compiled(self) -> bool:
        return self._compile

    def __call__(self, points: Optional[torch.Tensor] = None, func: Optional[Callable] = None,
                 point_data: Optional[Dict[str, torch.Tensor]] = None,
                 scalar_data: Optional[Dict[str, torch.Tensor]] = None, batch_size: int = 1) -> torch.Tensor:
        point_data = {} if point_data is None else dict(point_data)
        scalar_data = {} if scalar_data is None else dict(scalar_data)
        if points is None:
            points = self._points
        else:
            assert points.shape[1] == self.dimension, \
                f"the dimension of the points should be {self.dimension}, but got {points.shape[1]}"
        for key, value in point_data.items():
            assert value.shape[0] == points.shape[0], f"the shape of {key} should be [n_point, ...], but got {value.shape}"
        fn = self.forward if func is None else func
        match = self._match(fn, point_data, scalar_data) if points.device.type == "cuda" else None
        if match is not None:
            f = None if match.key is None else point_data[match.key]
            needs_grad = torch.is_grad_enabled() and (points.requires_grad or (f is not None and f.requires_grad))
            if not needs_grad:
                return self._native_linear(points, f, match.const)
        point_data["x"] = points
        return self._torch_linear(fn, points, point_data, scalar_data, batch_size)

    def _match(self, fn, point_data, scalar_data):
        """match_linear, cached per (integrand, its free state, data layout)."""
        layout = (tuple(sorted((k, tuple(v.shape[1:])) for k, v in point_data.items())), tuple(sorted(scalar_data)))
        cache = self.__dict__.setdefault("_match_cache", MatchCache())
        return cache.get(fn, layout, lambda: match_linear(fn, point_data, scalar_data))

    def _linear_args(self, points: torch.Tensor, f: Optional[torch.Tensor], const: float):
        """ctypes arguments of tm_assemble_linear(_noalloc) for a device call:
        (mesh view, form descriptor, constant or None, incidence view, the
        tensors the views point into).  ``points`` and ``f`` must be
        contiguous device tensors of one dtype."""
        dev = points.device
        tab, tab_dev = self._tables(dev)
        inc_ptr, inc = (t.to(dev) for t in self._incidence)
        cap = nat.reduce_cap(points.dtype, 1)
        plan = self._cached(f"nodeplan{cap}", dev,
                            lambda: torch.from_numpy(nat.reduce_plan(inc_ptr.cpu().numpy(), cap)).to(dev))
        mesh = self._mesh_view(points)
        form = nat.FormDesc(0, _n_quad(self.etype, self.quadrature_order), tab.ctypes.data, tab_dev.data_ptr(),
                            None if f is None else f.data_ptr(), 0.0, 0.0)
        iv = nat.IncidenceView(inc_ptr.data_ptr(), inc.data_ptr(), plan.numel() - 1, plan.data_ptr())
        cval = None if f is not None else ctypes.c_double(const)
        return mesh, form, cval, iv, (tab, tab_dev, inc_ptr, inc, plan, points, f, self._conn32(dev))

    def _native_linear(self, points, f, const) -> torch.Tensor:
        nat.require_device(points, "points")
        L = nat.lib()
        dev = points.device
        points = points.detach().contiguous()
        conn = self._conn32(dev)
        E, B = conn.shape
        if E == 0:
            return torch.zeros(points.shape[0], dtype=points.dtype, device=dev)
        if f is not None:
            nat.require_device(f, "point data")
            f = f.detach().to(points.dtype).contiguous()
        mesh, form, cval, iv, _keep = self._linear_args(points, f, const)
        slot = torch.empty(E * B, dtype=points.dtype, device=dev)
        out = torch.empty(points.shape[0], dtype=points.dtype, device=dev)
        nat.check(L.tm_assemble_linear(ctypes.byref(mesh), ctypes.byref(form),
                                       None if cval is None else ctypes.byref(cval), ctypes.byref(iv),
                                       slot.data_ptr(), out.data_ptr(), nat.stream_of(dev)),
                  "tm_assemble_linear")
        return out

    def _torch_linear(self, fn, points, point_data, scalar_data, batch_size) -> torch.Tensor:
        etype = self.etype
        conn = self.elements[etype].to(points.device)
        params = list(inspect.signature(fn).parameters)
        rules = [
            (lambda x: x == "v", (None, 0, 0)),
            (lambda x: x == "gradv", (0, 0, 0)),
            (lambda x: x in scalar_data, (None, None, None)),
            (lambda x: x in point_data, (0, 0, None)),
            (lambda x: x in {"grad" + k for k in point_data}, (0, 0, None)),
        ]
        dims = []
        for key in params:
            for cond, d in rules:
                if cond(key):
                    dims.append(d)
                    break
            else:
                raise ValueError(f"{key} is not supported, please use  `v`, `gradv` or more keys provided by "
                                 f"point_data, element_data or scalar_data")
        e_dims = tuple(d[0] for d in dims)
        q_dims = tuple(d[1] for d in dims)
        v_dims = tuple(d[2] for d in dims)
        element_parallel = not all(d is None for d in e_dims)
        inner = torch.vmap(torch.vmap(fn, in_dims=v_dims), in_dims=q_dims)
        pfn = torch.vmap(inner, in_dims=e_dims) if element_parallel else inner
        nq = _n_quad(etype, self.quadrature_order)
        if batch_size in (-1, None):
            n_batch, bs = 1, nq
        else:
            bs, n_batch = batch_size, math.ceil(nq / batch_size)
        ele_pd = {k: v[conn] for k, v in point_data.items()}
        total = None
        for i in range(n_batch):
            phi, grad, jxw = shape_grad_jxw(points, conn, etype, self.quadrature_order, i * bs, bs)
            args = []
            for key in params:
                if key == "v":
                    args.append(phi)
                elif key == "gradv":
                    args.append(grad)
                elif key in ele_pd:
                    args.append(torch.einsum("eb...,qb->eq...", ele_pd[key], phi))
                elif key.startswith("grad") and key[4:] in ele_pd:
                    args.append(torch.einsum("eb...,eqbd->eq...d", ele_pd[key[4:]], grad))
                elif key in scalar_data:
                    args.append(scalar_data[key])
            out = pfn(*args)
            out = torch.einsum("eqb...,eq->eb...", out, jxw) if element_parallel else \
                torch.einsum("qb...,eq->eb...", out, jxw)
            total = out if total is None else total + out
        flat = total.reshape(conn.numel(), *total.shape[2:])
        res = torch.zeros(self.n_points, *total.shape[2:], dtype=flat.dtype, device=flat.device)
        return res.index_add(0, conn.reshape(-1), flat).flatten()

    def __repr__(self):
        return (f"{self.__class__.__name__}(element_types={self.element_types}, "
                f"n_element={self.elements[self.etype].shape[0]}, n_point={self.n_points}, n_dim={self.dimension})")
