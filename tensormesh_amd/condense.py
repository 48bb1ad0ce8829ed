"""Static condensation of Dirichlet degrees of freedom.

Mirrors ``Condenser`` (tensormesh/operator/condense.py:26-443): same
constructor, ``__call__(K, rhs) -> (K_ii, f_i)``, ``condense_rhs``,
``recover``, ``restrict``, ``prolong`` and ``update_dirichlet``, the same
inner/outer edge masks in the same (order-preserving) entry order, and the
same layout-hash check.  The layout is computed once per pattern; the
per-call work -- gathering the K_ii values and forming
f_i - K_io u_o -- runs in one HIP kernel when the matrix lives on the device
(``tm_condense``: a coalesced gather of the kept entries and one thread per
inner row for the right-hand side).
"""

from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import torch
import torch.nn as nn

from . import _native as nat
from .sparse import SparseMatrix


class Condenser(nn.Module):
    def __init__(self, dirichlet_mask: torch.Tensor, dirichlet_value: Optional[torch.Tensor] = None):
        super().__init__()
        assert dirichlet_mask.dtype == torch.bool, f"dirichlet_mask must be a bool tensor, got {dirichlet_mask.dtype}"
        assert dirichlet_mask.ndim == 1, f"dirichlet_mask must be 1D, got shape {tuple(dirichlet_mask.shape)}"
        self.register_buffer("dirichlet_mask", dirichlet_mask)
        self.register_buffer("dirichlet_value", self._normalize_value(dirichlet_value))
        for name in ("inner_row", "inner_col", "ou2in_row", "ou2in_col", "is_inner_edge", "is_ou2in_edge",
                     "is_inner_dof", "is_outer_dof", "inner_idx", "ou2in_idx", "inner_row_ptr", "ou2in_row_ptr"):
            self.register_buffer(name, None, persistent=False)
        self.inner_shape = None
        self.ou2in_shape = None
        self.n_inner_dof = None
        self.n_outer_dof = None
        self.n_dof = None
        self.layout_hash = None
        self.K_ou2in = None
        self._inner_hash = None

    def _normalize_value(self, value: Optional[torch.Tensor]) -> torch.Tensor:
        mask = self.dirichlet_mask
        n_outer = int(mask.sum())
        if value is None:
            return torch.zeros(n_outer, device=mask.device)
        assert value.ndim == 1, f"dirichlet_value must be 1D, got shape {tuple(value.shape)}"
        if value.shape[0] == mask.shape[0]:
            return value[mask]
        assert value.shape[0] == n_outer, \
            f"dirichlet_value must have length n_dof ({mask.shape[0]}) or n_outer_dof ({n_outer}), got {value.shape[0]}"
        return value

    def _compute_layout(self, matrix: SparseMatrix):
        """condense.py:182-220: edge masks, local numbering, entry maps."""
        row, col = matrix.row, matrix.col
        n_dof = matrix.shape[0]
        mask = self.dirichlet_mask.to(row.device)
        inner, outer = ~mask, mask
        is_inner_edge = inner[row] & inner[col]
        is_ou2in_edge = inner[row] & outer[col]
        n_in, n_out = int(inner.sum()), int(outer.sum())
        local = torch.full((n_dof,), -1, dtype=torch.long, device=row.device)
        local[inner] = torch.arange(n_in, device=row.device)
        local[outer] = torch.arange(n_out, device=row.device)
        self.inner_idx = torch.nonzero(is_inner_edge).flatten()
        self.ou2in_idx = torch.nonzero(is_ou2in_edge).flatten()
        self.inner_row = local[row[self.inner_idx]]
        self.inner_col = local[col[self.inner_idx]]
        self.ou2in_row = local[row[self.ou2in_idx]]
        self.ou2in_col = local[col[self.ou2in_idx]]
        self.is_inner_edge, self.is_ou2in_edge = is_inner_edge, is_ou2in_edge
        self.is_inner_dof, self.is_outer_dof = inner, outer
        self.inner_shape = (n_in, n_in)
        self.ou2in_shape = (n_in, n_out)
        self.n_inner_dof, self.n_outer_dof, self.n_dof = n_in, n_out, n_dof
        self.layout_hash = matrix.layout_hash
        # CSR row pointers of the two sub-blocks (entries stay in row order)
        self.inner_row_ptr = self._row_ptr(self.inner_row, n_in)
        self.ou2in_row_ptr = self._row_ptr(self.ou2in_row, n_in)
        self._inner_hash = None

    @staticmethod
    def _row_ptr(rows: torch.Tensor, n: int) -> torch.Tensor:
        counts = torch.bincount(rows, minlength=n)
        rp = torch.zeros(n + 1, dtype=torch.int32, device=rows.device)
        rp[1:] = torch.cumsum(counts, 0).to(torch.int32)
        return rp

    def update_dirichlet(self, dirichlet_value: torch.Tensor):
        self.dirichlet_value = self._normalize_value(dirichlet_value)

    def _native_ok(self, matrix: SparseMatrix, rhs: torch.Tensor) -> bool:
        return (matrix.device.type == "cuda" and rhs.dim() == 1 and not matrix.values.requires_grad
                and not rhs.requires_grad and not self.dirichlet_value.requires_grad)

    def __call__(self, matrix: SparseMatrix, rhs: Optional[torch.Tensor] = None) -> Tuple[SparseMatrix, torch.Tensor]:
        if rhs is None:
            rhs = torch.zeros(matrix.shape[0], device=matrix.device)
        if self.inner_row is None:
            self._compute_layout(matrix)
        assert matrix.shape[0] == self.n_dof, f"the shape of matrix must be [{self.n_dof}, {self.n_dof}], but got {matrix.shape}"
        assert matrix.shape[1] == self.n_dof, f"the shape of matrix must be [{self.n_dof}, {self.n_dof}], but got {matrix.shape}"
        assert matrix.has_same_layout(self.layout_hash), \
            "the layout of the matrix is changed, please recompute the condensed matrix"
        assert rhs.shape[0] == self.n_dof, f"the shape of rhs must be [{self.n_dof}, ...], but got {rhs.shape}"
        dtype, dev = matrix.values.dtype, matrix.values.device
        self.dirichlet_value = self.dirichlet_value.to(dtype=dtype, device=dev)
        rhs = rhs.to(dtype=dtype, device=dev)
        if self._native_ok(matrix, rhs):
            kii_vals, f_i = self._native(matrix, rhs)
        else:
            kii_vals = matrix.values[self.inner_idx]
            kio = matrix.values[self.ou2in_idx]
            minus = torch.zeros(self.n_inner_dof, dtype=dtype, device=dev).index_add(
                0, self.ou2in_row, kio * self.dirichlet_value[self.ou2in_col])
            for _ in range(rhs.dim() - 1):
                minus = minus.unsqueeze(-1)
            f_i = rhs[self.is_inner_dof] - minus
        K_inner = SparseMatrix(kii_vals, self.inner_row, self.inner_col, self.inner_shape,
                               layout_hash=self._inner_hash, row_ptr=self.inner_row_ptr)
        if self._inner_hash is None:
            self._inner_hash = K_inner.layout_hash
        self.K_ou2in = SparseMatrix(matrix.values[self.ou2in_idx], self.ou2in_row, self.ou2in_col,
                                    self.ou2in_shape, row_ptr=self.ou2in_row_ptr)
        return K_inner, f_i

    def _native(self, matrix: SparseMatrix, rhs: torch.Tensor):
        L = nat.lib()
        dev = matrix.device
        vals = matrix.values.detach().contiguous()
        if not hasattr(self, "_cidx") or self._cidx[0].device != dev:
            inner_pos = torch.full((self.n_dof,), -1, dtype=torch.int32, device=dev)
            inner_pos[self.is_inner_dof] = torch.arange(self.n_inner_dof, dtype=torch.int32, device=dev)
            self._cidx = (self.inner_idx.to(torch.int32), self.inner_row_ptr, self.ou2in_idx.to(torch.int32),
                          self.ou2in_row_ptr, self.ou2in_col.to(torch.int32),
                          torch.nonzero(self.is_inner_dof).flatten().to(torch.int32))
        inner_idx, inner_rp, ou_idx, ou_rp, ou_col, inner_dofs = self._cidx
        kii = torch.empty(inner_idx.numel(), dtype=vals.dtype, device=dev)
        f_i = torch.empty(self.n_inner_dof, dtype=vals.dtype, device=dev)
        dval = self.dirichlet_value.to(vals.dtype).contiguous()
        nat.check(L.tm_condense(nat.dtype_code(vals.dtype), self.n_inner_dof, inner_idx.numel(), vals.data_ptr(),
                                inner_rp.data_ptr(), inner_idx.data_ptr(), ou_rp.data_ptr(), ou_idx.data_ptr(),
                                ou_col.data_ptr(), inner_dofs.data_ptr(), rhs.contiguous().data_ptr(),
                                dval.data_ptr(), kii.data_ptr(), f_i.data_ptr(), nat.stream_of(dev)),
                  "tm_condense")
        return kii, f_i

    def condense_rhs(self, rhs: torch.Tensor) -> torch.Tensor:
        assert self.K_ou2in is not None, \
            "call the Condenser on a matrix first; condense_rhs reuses the cached K_io block"
        self.dirichlet_value = self.dirichlet_value.type(rhs.dtype).to(rhs.device)
        rhs = rhs.type(self.K_ou2in.edata.dtype).to(self.K_ou2in.edata.device)
        minus = self.K_ou2in @ self.dirichlet_value
        for _ in range(rhs.dim() - 1):
            minus = minus.unsqueeze(-1)
        return rhs[self.is_inner_dof] - minus

    def recover(self, u: torch.Tensor) -> torch.Tensor:
        assert u.shape[0] == self.n_inner_dof, f"the shape of u must be [{self.n_inner_dof}, ...], but got {u.shape}"
        shape = list(u.shape)
        shape[0] = self.n_dof
        u_full = torch.zeros(shape, dtype=u.dtype, device=u.device)
        u_full[self.is_inner_dof.to(u.device)] += u
        bv = self.dirichlet_value.to(dtype=u.dtype, device=u.device)
        for _ in range(u.dim() - 1):
            bv = bv.unsqueeze(-1)
        u_full[self.is_outer_dof.to(u.device)] += bv
        return u_full

    def restrict(self, f: torch.Tensor) -> torch.Tensor:
        assert f.shape[0] == self.dirichlet_mask.shape[0], \
            f"the shape of f must be [{self.dirichlet_mask.shape[0]}, ...], but got {f.shape}"
        return f[~self.dirichlet_mask.to(f.device)]

    def prolong(self, f_inner: torch.Tensor) -> torch.Tensor:
        n_inner = int((~self.dirichlet_mask).sum())
        assert f_inner.shape[0] == n_inner, f"the shape of f_inner must be [{n_inner}, ...], but got {f_inner.shape}"
        shape = list(f_inner.shape)
        shape[0] = self.dirichlet_mask.shape[0]
        f_full = torch.zeros(shape, dtype=f_inner.dtype, device=f_inner.device)
        f_full[~self.dirichlet_mask.to(f_inner.device)] = f_inner
        return f_full
