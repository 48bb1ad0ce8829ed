"""Per-quadrature-point helpers used by the built-in forms' torch ``forward``.

Restated from tensormesh/functional/elasticity.py:132-284 and
tensormesh/functional/ops.py:8-155 (Voigt strain-displacement rows and the
isotropic Voigt stiffness).
"""

from __future__ import annotations

import torch


def voigt_shape_grad(gradu: torch.Tensor) -> torch.Tensor:
    """[d] gradient -> Voigt rows [3, 2] (xx, yy, xy) or [6, 3] (xx, yy, zz, yz, xz, xy)."""
    assert gradu.dim() == 1, f"gradu should be a 1D tensor of shape [dim], but got shape {gradu.shape}"
    dim = gradu.shape[-1]
    assert dim in (2, 3), f"dimension is only supported for 2, 3 for voigt shape grad, but got {dim}"
    z = torch.zeros_like(gradu[0])
    if dim == 2:
        g0, g1 = gradu[0], gradu[1]
        return torch.stack([torch.stack([g0, z]), torch.stack([z, g1]), torch.stack([g1, g0])])
    g0, g1, g2 = gradu[0], gradu[1], gradu[2]
    return torch.stack([torch.stack([g0, z, z]), torch.stack([z, g1, z]), torch.stack([z, z, g2]),
                        torch.stack([z, g2, g1]), torch.stack([g2, z, g0]), torch.stack([g1, g0, z])])


def voigt_stiffness(E, nu, dim: int = 2) -> torch.Tensor:
    """Isotropic Voigt stiffness with lambda = E nu / ((1+nu)(1-2nu)), mu = E / (2(1+nu))."""
    if dim not in (2, 3):
        raise Exception(f"dim must be 2 or 3, but got {dim}")
    mu = torch.as_tensor(E / (2.0 * (1.0 + nu)), dtype=torch.float64)
    lam = torch.as_tensor(E * nu / ((1 + nu) * (1 - 2 * nu)), dtype=torch.float64)
    n = 3 if dim == 2 else 6
    C = torch.zeros(n, n, dtype=torch.float64)
    C[:dim, :dim] = lam
    idx = torch.arange(dim)
    C[idx, idx] = lam + 2.0 * mu
    idx2 = torch.arange(dim, n)
    C[idx2, idx2] = mu
    return C
