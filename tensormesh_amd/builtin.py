"""Built-in forms (tensormesh/assemble/builtin.py:24-333, 730-813).

Each class keeps the reference's torch ``forward`` / ``element_energy`` so
that its semantics are visible and the generic path can run it; the
assembler recognises the integrand and runs the HIP kernels instead
(``tensormesh_amd.forms``).
"""

from __future__ import annotations

import torch

from .assemble import ElementAssembler, NodeAssembler
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
    and sums them by block rows (tm_assemble_neohookean_rows).
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
        return super().energy(point_data={"displacement": u})

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
