"""Sparse mat-vec and a Jacobi-preconditioned CG for SparseMatrix.

The reference inherits ``@`` and ``solve`` from ``torch_sla.SparseTensor``
(not vendored); its in-tree torch fallback is the Jacobi-CG of
tensormesh/sparse/solve/torch_solve.py:54-128 with the adjoint
``dL/dA_ij = -lambda_i u_j`` of :262-272.  Both are restated here with the
same semantics.
"""

from __future__ import annotations

from typing import Optional

import torch


def _spmv_raw(values: torch.Tensor, row: torch.Tensor, col: torch.Tensor, n_rows: int,
              x: torch.Tensor) -> torch.Tensor:
    prod = values.reshape(-1, *([1] * (x.dim() - 1))) * x[col]
    out = torch.zeros((n_rows,) + tuple(x.shape[1:]), dtype=prod.dtype, device=x.device)
    return out.index_add(0, row, prod)


def spmv(A, x: torch.Tensor) -> torch.Tensor:
    """``A @ x`` for a vector or a stack of column vectors; differentiable in
    ``A.values`` and ``x``."""
    assert x.shape[0] == A.shape[1], f"shape mismatch {A.shape} @ {tuple(x.shape)}"
    return _spmv_raw(A.values, A.row_indices, A.col_indices, A.shape[0], x.to(A.values.dtype))


def cg(A, b: torch.Tensor, x0: Optional[torch.Tensor] = None, rtol: float = 1e-10, atol: float = 0.0,
       max_iter: int = 100000, check_every: int = 25):
    """Jacobi-preconditioned conjugate gradients (torch_solve.py:54-128).

    Returns (x, iterations, relative residual).  Convergence is tested every
    ``check_every`` iterations to keep host synchronisation rare.
    """
    vals = A.values.detach()
    row, col = A.row_indices, A.col_indices
    n = A.shape[0]
    diag = torch.zeros(n, dtype=vals.dtype, device=vals.device).index_add(
        0, row, torch.where(row == col, vals, torch.zeros_like(vals)))
    dinv = torch.where(diag != 0, 1.0 / diag, torch.ones_like(diag))
    b = b.detach().to(vals.dtype)
    x = torch.zeros_like(b) if x0 is None else x0.detach().clone().to(vals.dtype)
    r = b - _spmv_raw(vals, row, col, n, x) if x0 is not None else b.clone()
    z = dinv * r
    p = z.clone()
    rz = torch.dot(r, z)
    bnorm = torch.linalg.vector_norm(b).item()
    tol = max(rtol * bnorm, atol)
    it = 0
    rnorm = torch.linalg.vector_norm(r).item()
    if rnorm <= tol:
        return x, 0, rnorm / max(bnorm, 1e-300)
    while it < max_iter:
        Ap = _spmv_raw(vals, row, col, n, p)
        alpha = rz / torch.dot(p, Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        it += 1
        if it % check_every == 0:
            rnorm = torch.linalg.vector_norm(r).item()
            if rnorm <= tol:
                break
        z = dinv * r
        rz_new = torch.dot(r, z)
        p = z + (rz_new / rz) * p
        rz = rz_new
    rnorm = torch.linalg.vector_norm(b - _spmv_raw(vals, row, col, n, x)).item()
    return x, it, rnorm / max(bnorm, 1e-300)


class _Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, b, A, kwargs):
        x, it, rel = cg(A, b, **kwargs)
        ctx.A, ctx.kwargs = A, kwargs
        ctx.save_for_backward(x)
        ctx.info = (it, rel)
        return x

    @staticmethod
    def backward(ctx, gx):
        (x,) = ctx.saved_tensors
        A = ctx.A
        # symmetric A: A^T lam = gx is the same system (torch_solve.py:262-272)
        lam, _, _ = cg(A, gx, **ctx.kwargs)
        gvals = -lam[A.row_indices] * x[A.col_indices]
        return gvals, lam, None, None


def solve(A, b: torch.Tensor, **kwargs) -> torch.Tensor:
    """x = A^-1 b for symmetric positive definite A, differentiable in
    ``A.values`` and ``b``."""
    assert A.shape[0] == A.shape[1] == b.shape[0]
    kwargs.pop("verbose", None)
    return _Solve.apply(A.values, b, A, kwargs)
