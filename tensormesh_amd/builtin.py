"""Built-in forms (tensormesh/assemble/builtin.py:24-561, 730-813).

Each class keeps the reference's torch ``forward`` / ``element_energy`` so
that its semantics are visible and the generic path can run it; the
assembler recognises the integrand and runs the HIP kernels instead
(``tensormesh_amd.forms``).
"""

from __future__ import annotations

import ctypes
import math
from typing import Optional

import torch
from torch.autograd.function import once_differentiable

from . import _native as nat
from .assemble import ElementAssembler, NodeAssembler, _n_quad
from .functional import voigt_shape_grad, voigt_stiffness


class LaplaceElementAssembler(ElementAssembler):
    """K_ij = int grad(N_i) . grad(N_j) (builtin.py:24-62)."""

    def forward(self, gradu, gradv):
        return gradu @ gradv


class DiffusionElementAssembler(ElementAssembler):
    """K_ij = int kappa grad(N_i) . grad(N_j) with ``element_data={'kappa': [E]}``
    or ``point_data={'kappa': [N]}`` -- the SIMP / inverse-design form of
    examples/inverse_design/thermal_topology.py:48-52 as a named class."""

    def forward(self, gradu, gradv, kappa):
        return kappa * (gradu @ gradv)


class MassElementAssembler(ElementAssembler):
    """M_ij = int N_i N_j (builtin.py:64-100)."""

    def forward(self, u, v):
        return u * v


class LinearElasticityElementAssembler(ElementAssembler):
    """Ba^T C Bb with isotropic C(E, nu) (builtin.py:102-217)."""

    def __post_init__(self, E=1.0, nu=0.3):
        self.E = E
        self.nu = nu

    def forward(self, gradu, gradv):
        dim = gradu.shape[0]
        Ba = voigt_shape_grad(gradu)
        Bb = voigt_shape_grad(gradv)
        C = voigt_stiffness(self.E, self.nu, dim).to(dtype=gradu.dtype, device=gradu.device)
        return Ba.T @ C @ Bb

    def element_energy(self, graddisplacement):
        grad_u = graddisplacement
        eps = 0.5 * (grad_u + grad_u.transpose(-1, -2))
        mu = self.E / (2 * (1 + self.nu))
        lam = (self.E * self.nu) / ((1 + self.nu) * (1 - 2 * self.nu))
        tr = eps.diagonal(dim1=-2, dim2=-1).sum(-1)
        return 0.5 * lam * tr ** 2 + mu * (eps * eps).sum(dim=(-2, -1))


class NeoHookeanModel(ElementAssembler):
    """psi = mu/2 (I1 - d) - mu ln J + lam/2 (ln J)^2 (builtin.py:219-333).

    ``nh_path = "slots"`` (default) writes every element's block rows in
    element order and sums them with the segment reduction
    (tm_assemble_neohookean); ``"rows"`` writes them in node-incidence order
    and sums them by block rows (tm_assemble_neohookean_rows); ``"auto"``
    times both once per device and pattern and keeps the faster
    (tensormesh_amd.tune).
    ``use_mfma = True`` (float32) contracts the tangent on the matrix cores
    (v_mfma_f32_32x32x2f32) instead of the closed-form VALU loop."""

    use_mfma: bool = False
    nh_path: str = "slots"

    def __post_init__(self, E=1.0, nu=0.3):
        self.E = E
        self.nu = nu
        self.mu = E / (2 * (1 + nu))
        self.lam = E * nu / ((1 + nu) * (1 - 2 * nu))

    def element_energy(self, graddisplacement):
        grad_u = graddisplacement
        dim = grad_u.shape[-1]
        F = torch.eye(dim, device=grad_u.device, dtype=grad_u.dtype) + grad_u
        J = torch.linalg.det(F)
        I1 = (F * F).sum()
        log_J = torch.log(J)
        return (self.mu / 2) * (I1 - dim) - self.mu * log_J + (self.lam / 2) * (log_J ** 2)

    def energy(self, u):
        """Total strain energy (builtin.py:316-333).  On the ROCm device it
        runs on tm_neohookean_energy, differentiable in ``u`` (backward is the
        HIP residual, assembled in the same launch when ``u`` requires a
        gradient); on the CPU, or with ``reference_path``, it runs the
        reference's generic energy path."""
        if (u.device.type == "cuda" and not self.reference_path and not self._points.requires_grad
                and u.dim() == 2 and tuple(u.shape) == tuple(self._points.shape)):
            return _NHEnergy.apply(u, self)
        return super().energy(point_data={"displacement": u})

    def _native_energy(self, u: torch.Tensor, want_residual: bool, points: Optional[torch.Tensor] = None):
        """(energy [] in the mesh dtype, residual [N*D] or None) on tm_neohookean_energy."""
        points = self._points if points is None else points
        nat.require_device(points, "points")
        L = nat.lib()
        dev = points.device
        dtype = points.dtype
        points = points.detach().contiguous()
        u = u.detach().to(dtype).contiguous()
        conn = self._conn32(dev)
        E, B = conn.shape
        D = self.dimension
        pat = self._pattern_on(dev)
        tab, tab_dev = self._tables(dev)
        mesh = self._mesh_view(points)
        form = nat.FormDesc(0, _n_quad(self.etype, self.quadrature_order), tab.ctypes.data, tab_dev.data_ptr(),
                            None, float(self.E), float(self.nu), 0)
        form.epw = self._launch_epw("neohookean_energy", dev, dtype, str(int(want_residual)),
                                    lambda e: self._native_energy(u, want_residual, points), (1, 2, 4))
        ews = torch.empty(E + nat.J2_ENERGY_PARTS, dtype=torch.float64, device=dev)
        energy = torch.empty((), dtype=dtype, device=dev)
        res = rslot = None
        iv = None
        if want_residual:
            node_plan = pat.plan(nat.reduce_cap(dtype, D), node=True)
            iv = nat.IncidenceView(pat.inc_ptr.data_ptr(), pat.inc.data_ptr(), node_plan.numel() - 1,
                                   node_plan.data_ptr())
            rslot = torch.empty(E * B * D, dtype=dtype, device=dev)
            res = torch.empty(self.n_points * D, dtype=dtype, device=dev)
        p = lambda t: None if t is None else t.data_ptr()
        nat.check(L.tm_neohookean_energy(ctypes.byref(mesh), ctypes.byref(form), u.data_ptr(), ews.data_ptr(),
                                         energy.data_ptr(), None if iv is None else ctypes.byref(iv), p(rslot),
                                         p(res), nat.stream_of(dev)), "tm_neohookean_energy")
        return energy, res

    # The reference obtains these by autograd of ``energy`` (element_assembler.py:447-659);
    # here they are assembled directly on the HIP kernels.
    def residual(self, u, points=None):
        """dE/du, flattened [N*D] (node-major, as NodeAssembler vector output)."""
        return self._native_hyperelastic(u, points, self.E, self.nu, True, False)[0]

    def tangent(self, u, points=None):
        """d2E/du2 as a block SparseMatrix [N*D, N*D] in the reference's block layout."""
        return self._native_hyperelastic(u, points, self.E, self.nu, False, True)[1]

    def residual_and_tangent(self, u, points=None):
        return self._native_hyperelastic(u, points, self.E, self.nu, True, True)


class _NHEnergy(torch.autograd.Function):
    """Neo-Hookean energy on tm_neohookean_energy; its backward is the HIP
    residual (the reference's ``energy(u).backward()``, as in
    examples/solid/hyperelastic_beam/hyperelastic_beam.py:133-175).  When
    ``u`` requires a gradient the forward launch assembles the residual beside
    the energy and backward only scales it.  Second derivatives come from
    ``NeoHookeanModel.tangent``, not from double backward."""

    @staticmethod
    def forward(ctx, u, model):
        want_grad = ctx.needs_input_grad[0]
        energy, res = model._native_energy(u, want_residual=want_grad)
        ctx.save_for_backward(res if want_grad else None)
        ctx.u_shape, ctx.u_dtype = u.shape, u.dtype
        return energy

    @staticmethod
    @once_differentiable
    def backward(ctx, g):
        (res,) = ctx.saved_tensors
        return (g * res).reshape(ctx.u_shape).to(ctx.u_dtype), None


class _J2Energy(torch.autograd.Function):
    """Energy on tm_assemble_j2; its backward is the HIP residual (the
    reference's ``energy().backward()``).  When the displacement requires a
    gradient, the forward launch assembles the residual beside the energy (one
    element pass instead of two) and backward only scales it.  Second
    derivatives come from ``J2Plasticity.tangent``, not from double backward."""

    @staticmethod
    def forward(ctx, u, eps_p, alpha, model, points):
        want_grad = ctx.needs_input_grad[0]
        energy, res, _ = model._native_j2(u, eps_p, alpha, points, want_energy=True, want_residual=want_grad)
        ctx.save_for_backward(res if want_grad else None)
        ctx.u_shape, ctx.u_dtype = u.shape, u.dtype
        return energy

    @staticmethod
    @once_differentiable
    def backward(ctx, g):
        (res,) = ctx.saved_tensors
        return (g * res).reshape(ctx.u_shape).to(ctx.u_dtype), None, None, None, None


class J2Plasticity(ElementAssembler):
    """J2 (von Mises) plasticity with linear isotropic hardening
    (builtin.py:335-561): return-mapping incremental potential per
    quadrature point, history ``{etype: {'eps_p': [E,Q,3,3], 'alpha': [E,Q]}}``.

    ``energy(point_data={'displacement': u}, element_data={'eps_p_n': ...,
    'alpha_n': ...})`` runs on tm_assemble_j2 when the data is on the device
    (differentiable in ``u``: backward is the HIP residual); other calls use
    the generic torch path.  ``residual`` / ``tangent`` assemble dE/du and the
    consistent Hessian directly (the reference has only autograd of the
    energy), and ``update_state`` is the radial return on tm_j2_update_state.
    """

    def __post_init__(self, material=None, E=200e9, nu=0.3, sig0=250e6, H=1e9):
        if material is not None:
            self.E, self.nu, self.sig0, self.H = material.E, material.nu, material.sigma_y, material.H
        else:
            self.E, self.nu, self.sig0, self.H = E, nu, sig0, H
        self.mu = self.E / (2 * (1 + self.nu))
        self.lam = (self.E * self.nu) / ((1 + self.nu) * (1 - 2 * self.nu))
        self.bulk = self.lam + 2 / 3 * self.mu
        nq = _n_quad(self.etype, self.quadrature_order)
        n_elem = self.elements[self.etype].shape[0]
        self.history = {self.etype: {
            "eps_p": torch.zeros((n_elem, nq, 3, 3), device=self.device, dtype=self.dtype),
            "alpha": torch.zeros((n_elem, nq), device=self.device, dtype=self.dtype)}}

    def element_energy(self, graddisplacement, eps_p_n, alpha_n):
        """psi at one quadrature point (builtin.py:464-511)."""
        grad_u = graddisplacement
        dim = grad_u.shape[0]
        eps = 0.5 * (grad_u + grad_u.T)
        if dim == 2:
            eps = torch.nn.functional.pad(eps, (0, 1, 0, 1))
        eps_tr = eps - eps_p_n
        tr = eps_tr.diagonal(dim1=-2, dim2=-1).sum(-1)
        dev = eps_tr - (tr / 3.0) * torch.eye(3, device=grad_u.device, dtype=grad_u.dtype)
        nrm = torch.norm(dev)
        f_tr = 2 * self.mu * nrm - math.sqrt(2 / 3) * (self.sig0 + self.H * alpha_n)
        denom = 2 * self.mu + (2 / 3) * self.H
        dg = torch.clamp(f_tr, min=0.0) / denom
        return 0.5 * self.bulk * tr ** 2 + self.mu * nrm ** 2 - 0.5 * denom * dg ** 2

    # ---- energy (element_assembler.py:447-659) ----
    def energy(self, points=None, func=None, point_data=None, element_data=None, scalar_data=None,
               batch_size: int = -1):
        hist = self._history_args(element_data)
        u = (point_data or {}).get("displacement")
        native = (func is None and not self.reference_path and not scalar_data and u is not None
                  and set(point_data) == {"displacement"} and hist is not None
                  and u.device.type == "cuda" and (points is None or not points.requires_grad)
                  and not hist[0].requires_grad and not hist[1].requires_grad)
        if not native:
            return super().energy(points, func, point_data, element_data, scalar_data, batch_size)
        return _J2Energy.apply(u, hist[0], hist[1], self, points)

    def _history_args(self, element_data):
        """(eps_p_n, alpha_n) from element_data given as tensors or {etype: tensor}."""
        if not element_data or set(element_data) != {"eps_p_n", "alpha_n"}:
            return None
        out = []
        for key in ("eps_p_n", "alpha_n"):
            v = element_data[key]
            out.append(v[self.etype] if isinstance(v, dict) else v)
        return out

    def _state(self, eps_p, alpha):
        h = self.history[self.etype]
        return (h["eps_p"] if eps_p is None else eps_p), (h["alpha"] if alpha is None else alpha)

    def residual(self, u, eps_p=None, alpha=None, points=None):
        """dE/du [N*D] at the history (default: ``self.history``)."""
        return self._native_j2(u, *self._state(eps_p, alpha), points, want_residual=True)[1]

    def tangent(self, u, eps_p=None, alpha=None, points=None):
        """Consistent tangent d2E/du2 as a block SparseMatrix [N*D, N*D]."""
        return self._native_j2(u, *self._state(eps_p, alpha), points, want_tangent=True)[2]

    def residual_and_tangent(self, u, eps_p=None, alpha=None, points=None):
        _, r, k = self._native_j2(u, *self._state(eps_p, alpha), points, want_residual=True, want_tangent=True)
        return r, k

    def update_state(self, u_vec):
        """Radial return after a converged step (builtin.py:513-561)."""
        with torch.no_grad():
            h = self.history[self.etype]
            if u_vec.device.type != "cuda":
                h["eps_p"], h["alpha"] = self._update_state_torch(u_vec, h["eps_p"], h["alpha"])
                return
            points = self._points
            nat.require_device(points, "points")
            dtype = points.dtype
            for k in ("eps_p", "alpha"):
                if h[k].device != points.device or h[k].dtype != dtype or not h[k].is_contiguous():
                    h[k] = h[k].to(device=points.device, dtype=dtype).contiguous()
            u = u_vec.detach().to(dtype).contiguous()
            tab, tab_dev = self._tables(points.device)
            form = nat.FormDesc(0, _n_quad(self.etype, self.quadrature_order), tab.ctypes.data, tab_dev.data_ptr(),
                                None, 0.0, 0.0, 0)
            st = nat.J2State(h["eps_p"].data_ptr(), h["alpha"].data_ptr())
            nat.check(nat.lib().tm_j2_update_state(ctypes.byref(self._mesh_view(points)), ctypes.byref(form),
                                                   ctypes.byref(self._material()), u.data_ptr(), ctypes.byref(st),
                                                   nat.stream_of(points.device)), "tm_j2_update_state")

    def _update_state_torch(self, u_vec, eps_p, alpha):
        """Generic torch path of update_state for host tensors (builtin.py:513-561)."""
        from .assemble import shape_grad_jxw
        conn = self.elements[self.etype].to(u_vec.device)
        _, grad, _ = shape_grad_jxw(self._points.to(u_vec.device), conn, self.etype, self.quadrature_order)
        G = torch.einsum("bqkx,bku->bqux", grad, u_vec[conn])
        eps = torch.zeros(G.shape[:2] + (3, 3), device=u_vec.device, dtype=u_vec.dtype)
        d = G.shape[-1]
        eps[..., :d, :d] = 0.5 * (G + G.transpose(-1, -2))
        eps_tr = eps - eps_p
        tr = eps_tr.diagonal(dim1=-2, dim2=-1).sum(-1)
        dev = eps_tr - (tr[..., None, None] / 3.0) * torch.eye(3, device=u_vec.device, dtype=u_vec.dtype)
        nrm = torch.norm(dev, dim=(-2, -1))
        f_tr = 2 * self.mu * nrm - math.sqrt(2 / 3) * (self.sig0 + self.H * alpha)
        dg = torch.clamp(f_tr, min=0.0) / (2 * self.mu + (2 / 3) * self.H)
        safe = torch.where(nrm < 1e-12, torch.ones_like(nrm), nrm)
        dgm = dg * (f_tr > 0).to(dg.dtype)
        return eps_p + dgm[..., None, None] * dev / safe[..., None, None], alpha + math.sqrt(2 / 3) * dgm

    def _material(self):
        return nat.J2Material(float(self.E), float(self.nu), float(self.sig0), float(self.H))

    def _native_j2(self, u, eps_p, alpha, points=None, want_energy=False, want_residual=False,
                   want_tangent=False):
        """(energy [], residual [N*D], tangent SparseMatrix) on tm_assemble_j2; unrequested parts are None."""
        points = self._points if points is None else points
        nat.require_device(points, "points")
        nat.require_device(u, "displacement")
        L = nat.lib()
        dev, dtype = points.device, points.dtype
        points = points.detach().contiguous()
        u = u.detach().to(dtype).contiguous()
        assert u.shape == points.shape, f"displacement must be [{points.shape[0]}, {points.shape[1]}]"
        conn = self._conn32(dev)
        E, B = conn.shape
        D = self.dimension
        nq = _n_quad(self.etype, self.quadrature_order)
        eps_p = eps_p.detach().to(device=dev, dtype=dtype).contiguous()
        alpha = alpha.detach().to(device=dev, dtype=dtype).contiguous()
        assert tuple(eps_p.shape) == (E, nq, 3, 3), f"eps_p_n must be [{E}, {nq}, 3, 3], got {tuple(eps_p.shape)}"
        assert tuple(alpha.shape) == (E, nq), f"alpha_n must be [{E}, {nq}], got {tuple(alpha.shape)}"
        pat = self._pattern_on(dev)
        tab, tab_dev = self._tables(dev)
        sym = want_tangent and self._symmetric_blocks(True)
        form = nat.FormDesc(0, nq, tab.ctypes.data, tab_dev.data_ptr(), None, 0.0, 0.0,
                            nat.TM_FLAG_UPPER if sym else 0)
        form.epw = self._launch_epw("j2", dev, dtype, f"{int(sym)}|{int(want_energy)}|{int(want_residual)}|"
                                    f"{int(want_tangent)}",
                                    lambda e: self._native_j2(u, eps_p, alpha, points, want_energy=want_energy,
                                                              want_residual=want_residual,
                                                              want_tangent=want_tangent), (1, 4, 8))
        st = nat.J2State(eps_p.data_ptr(), alpha.data_ptr())
        pv = pat.view(dtype, D * D, symmetric=sym)
        node_plan = pat.plan(nat.reduce_cap(dtype, D), node=True)
        iv = nat.IncidenceView(pat.inc_ptr.data_ptr(), pat.inc.data_ptr(), node_plan.numel() - 1, node_plan.data_ptr())
        new = lambda n, want, dt=dtype: torch.empty(n, dtype=dt, device=dev) if want else None
        ews, energy = new(E + nat.J2_ENERGY_PARTS, want_energy, torch.float64), new((), want_energy)
        pairs = B * (B + 1) // 2 if sym else B * B      # packed upper blocks, or all
        tslot, tan = new(E * pairs * D * D, want_tangent), new(pat.nnz * D * D, want_tangent)
        rslot, res = new(E * B * D, want_residual), new(self.n_points * D, want_residual)
        nat.check(L.tm_assemble_j2(ctypes.byref(self._mesh_view(points)), ctypes.byref(form),
                                   ctypes.byref(self._material()), u.data_ptr(), ctypes.byref(st), nat.ptr(ews),
                                   nat.ptr(energy), ctypes.byref(pv), ctypes.byref(iv), nat.ptr(tslot),
                                   nat.ptr(tan), nat.ptr(rslot), nat.ptr(res), nat.stream_of(dev)),
                  "tm_assemble_j2")
        K = self._wrap_output(tan.reshape(pat.nnz, D, D), dev) if want_tangent else None
        return energy, res, K


def const_node_assembler(c=1):
    """NodeAssembler subclass for f_i = int c N_i (builtin.py:730-770)."""

    class ConstNodeAssembler(NodeAssembler):
        def __post_init__(self, c=c):
            self.c = c

        def forward(self, v):
            return self.c * v

    return ConstNodeAssembler


def func_node_assembler(f=lambda x: x):
    """NodeAssembler subclass for f_i = int f(x) N_i (builtin.py:772-813)."""

    class FuncNodeAssembler(NodeAssembler):
        def __post_init__(self, f=f):
            self.f = f

        def forward(self, x, v):
            return self.f(x) * v

    return FuncNodeAssembler


class SourceNodeAssembler(NodeAssembler):
    """b_i = int f_h N_i with nodal ``point_data={'f': [N]}`` (README.md:103-118)."""

    def forward(self, v, f):
        return f * v
