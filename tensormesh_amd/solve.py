"""Sparse mat-vec and Jacobi-preconditioned CG for SparseMatrix.

The reference inherits ``@`` and ``solve`` from ``torch_sla.SparseTensor``
(not vendored); its in-tree torch fallback is the Jacobi-CG of
tensormesh/sparse/solve/torch_solve.py:54-128 with the adjoint
``dL/dA_ij = -lambda_i u_j`` of :262-272.  Here:

* on the ROCm device the solve runs in libtmhip's batched PCG
  (``tm_pcg_init`` / ``tm_pcg_iterate``): iterations are launched in chunks,
  optionally captured in a HIP graph, and the host only polls the per-system
  convergence flags between chunks;
* on the CPU the same algorithm runs in torch (host logic for small
  problems and tests);
* ``solve`` is differentiable in the matrix values and the right-hand side
  for symmetric matrices (the adjoint system is the forward one).
"""

from __future__ import annotations

import ctypes
import warnings
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _native as nat

BLOCK_ROWS = 256


def _spmv_raw(values, row, col, n_rows, x):
    prod = values.reshape(-1, *([1] * (x.dim() - 1))) * x[col]
    out = torch.zeros((n_rows,) + tuple(x.shape[1:]), dtype=prod.dtype, device=x.device)
    return out.index_add(0, row, prod)


def spmv(A, x: torch.Tensor) -> torch.Tensor:
    """``A @ x``; differentiable in ``A.values`` and ``x``.  Uses the native
    CSR kernel for device matrices in CSR order when no gradient is needed."""
    assert x.shape[0] == A.shape[1], f"shape mismatch {A.shape} @ {tuple(x.shape)}"
    x = x.to(A.values.dtype)
    needs_grad = torch.is_grad_enabled() and (A.values.requires_grad or x.requires_grad)
    if (A.device.type == "cuda" and A.row_ptr is not None and x.dim() == 1 and not needs_grad
            and A.values.dtype in (torch.float32, torch.float64)):
        y = torch.empty(A.shape[0], dtype=A.values.dtype, device=A.device)
        xc = x.contiguous()
        nat.check(nat.lib().tm_spmv(nat.dtype_code(A.values.dtype), A.shape[0], A.row_ptr.data_ptr(),
                                    _col32(A).data_ptr(), A.values.contiguous().data_ptr(), xc.data_ptr(),
                                    y.data_ptr(), nat.stream_of(A.device)), "tm_spmv")
        return y
    return _spmv_raw(A.values, A.row_indices, A.col_indices, A.shape[0], x)


def _col32(A) -> torch.Tensor:
    c = getattr(A, "_col32", None)
    if c is None or c.device != A.device:
        c = A.col_indices.to(torch.int32).contiguous()
        A._col32 = c
    return c


# ------------------------------------------------------------ native PCG ---
class BatchedPCG:
    """Jacobi-PCG over a block-diagonal stack of CSR systems on the device.

    ``launches`` picks the iteration: 2 (default) the two-launch
    Chronopoulos-Gear iteration, 1 the same iteration in one launch per step
    (the SpMV recomputes the updated vector of every column it reads: four
    gathers per matrix entry instead of one, measured 2.5x slower per
    iteration on a 261k-row system) -- both for stacks whose largest system
    spans at most ``TWO_LAUNCH_MAX_BLOCKS`` workgroups (every workgroup
    re-sums its system's partials) and with bit-identical iterates -- or 5,
    the textbook PCG (used as well for larger systems).

    Parameters
    ----------
    row_ptr, col_idx : int32 device tensors of the stacked CSR (global ids)
    values : device tensor, float32 or float64
    sys_rows : row offsets of the systems, length n_sys + 1
    rtol : stop a system once |r| <= rtol |b|
    """

    launches: int = 2
    TWO_LAUNCH_MAX_BLOCKS = 2048

    def __init__(self, row_ptr: torch.Tensor, col_idx: torch.Tensor, values: torch.Tensor,
                 sys_rows: Sequence[int], rtol: float = 1e-10):
        self.dev = values.device
        self.dtype = values.dtype
        self.n = int(row_ptr.numel() - 1)
        self.row_ptr = row_ptr.to(torch.int32).contiguous()
        self.col_idx = col_idx.to(torch.int32).contiguous()
        self.values = values.detach().clone().contiguous()   # owned: graph replays read this buffer
        sys_rows = np.asarray(sys_rows, dtype=np.int64)
        self.n_sys = int(sys_rows.size - 1)
        blk_sys, blk0, blk1, sys_blk = [], [], [], [0]
        for s in range(self.n_sys):
            a, b = int(sys_rows[s]), int(sys_rows[s + 1])
            for r0 in range(a, b, BLOCK_ROWS):
                blk_sys.append(s)
                blk0.append(r0)
                blk1.append(min(b, r0 + BLOCK_ROWS))
            sys_blk.append(len(blk_sys))
        t = lambda a: torch.tensor(a, dtype=torch.int32, device=self.dev)
        self.blk_sys, self.blk_row0, self.blk_row1, self.sys_blk = t(blk_sys), t(blk0), t(blk1), t(sys_blk)
        self.n_blocks = len(blk_sys)
        max_blk = max((sys_blk[i + 1] - sys_blk[i] for i in range(self.n_sys)), default=0)
        small = max_blk <= self.TWO_LAUNCH_MAX_BLOCKS
        self.variant = {1: 2, 2: 1}.get(int(self.launches), 0) if small else 0   # tm_pcg_view.two_launch
        self.two = self.variant != 0
        self.stride = 16 if self.two else 8
        self._parity = 0
        e = lambda: torch.empty(self.n, dtype=self.dtype, device=self.dev)
        self.dinv, self.r, self.z, self.p, self.q, self.x, self.s = e(), e(), e(), e(), e(), e(), e()
        self.aux = torch.empty(2 * self.n if self.variant == 2 else 1, dtype=self.dtype, device=self.dev)
        n_part = {0: 3, 1: 5, 2: 7}[self.variant]
        self.partials = torch.zeros(self.n_blocks * n_part, dtype=torch.float64, device=self.dev)
        self.scalars = torch.zeros(self.n_sys * self.stride, dtype=torch.float64, device=self.dev)
        self.rtol = rtol
        self._graphs = {}
        self._last_iters = 0
        self.refresh_values(None)

    def refresh_values(self, values: Optional[torch.Tensor]):
        """New matrix values on the same pattern (copied into the owned buffer,
        so captured graphs stay valid); recomputes the Jacobi diagonal."""
        if values is not None:
            self.values.copy_(values.detach())
        nat.check(nat.lib().tm_jacobi(nat.dtype_code(self.dtype), self.n, self.row_ptr.data_ptr(),
                                      self.col_idx.data_ptr(), self.values.data_ptr(), self.dinv.data_ptr(),
                                      nat.stream_of(self.dev)), "tm_jacobi")

    def _view(self) -> nat.PcgView:
        v = nat.PcgView()
        v.dtype, v.reserved = nat.dtype_code(self.dtype), 0
        v.n_rows, v.n_sys, v.n_blocks = self.n, self.n_sys, self.n_blocks
        for name in ("row_ptr", "col_idx", "values", "dinv", "r", "z", "p", "q", "partials", "scalars",
                     "blk_sys", "blk_row0", "blk_row1", "sys_blk"):
            setattr(v, name, getattr(self, name).data_ptr())
        v.rtol = self.rtol
        if self.two:
            v.s, v.two_launch, v.parity = self.s.data_ptr(), self.variant, self._parity
            v.aux = self.aux.data_ptr()
        return v

    def residual(self, b: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
        """r = b - A x into ``self.r`` (device, this solver's precision) and the
        per-system relative residual |r_s| / |b_s| (float64 tensor)."""
        view = self._view()
        bb = b.to(self.dtype).contiguous()
        xx = x.to(self.dtype).contiguous()
        nat.check(nat.lib().tm_pcg_init(ctypes.byref(view), bb.data_ptr(), xx.data_ptr(),
                                        nat.stream_of(self.dev)), "tm_pcg_init")
        self._parity = 0
        sc = self.scalars.view(self.n_sys, self.stride)
        return torch.sqrt(sc[:, 4] / torch.clamp(sc[:, 5], min=1e-300))

    def _iterate(self, n):
        view = self._view()
        nat.check(nat.lib().tm_pcg_iterate(ctypes.byref(view), self.x.data_ptr(), n, nat.stream_of(self.dev)),
                  "tm_pcg_iterate")
        self._parity = (self._parity + n) & 1

    def solve(self, b: torch.Tensor, x0: Optional[torch.Tensor] = None, max_iter: int = 20000,
              chunk: Optional[int] = None, graph: bool = True):
        """Returns (x, iterations per system [n_sys], relative residual per system).

        ``chunk``: iterations per graph replay between convergence polls; None
        picks it from the previous solve's iteration count of this solver
        (tensormesh_amd.tune.pcg_chunk).  Converged systems stop on the device,
        so the chunk does not change the result."""
        if chunk is None:
            from .tune import pcg_chunk
            chunk = pcg_chunk(self._last_iters)
        view = self._view()
        if x0 is None:
            self.x.zero_()
        else:
            self.x.copy_(x0)
        bb = b.to(self.dtype).contiguous()
        nat.check(nat.lib().tm_pcg_init(ctypes.byref(view), bb.data_ptr(), self.x.data_ptr(),
                                        nat.stream_of(self.dev)), "tm_pcg_init")
        self._parity = 0
        if self.two:
            chunk += chunk & 1  # graphs replay with fixed parities: keep chunks even
        done = 0
        g = None
        if graph:
            g = self._graphs.get(chunk)
            if g is None:
                s = torch.cuda.Stream(device=self.dev)
                s.wait_stream(torch.cuda.current_stream(self.dev))
                g = torch.cuda.CUDAGraph()
                with torch.cuda.stream(s):
                    with torch.cuda.graph(g, stream=s):
                        self._iterate(chunk)
                torch.cuda.current_stream(self.dev).wait_stream(s)
                self._graphs[chunk] = g
                # capturing does not execute: start from the initial state again
                self._parity = 0
        while done < max_iter:
            if g is not None:
                g.replay()
            else:
                self._iterate(chunk)
            done += chunk
            sc = self.scalars.view(self.n_sys, self.stride)
            if not bool((sc[:, 6] != 0).any().item()):
                break
        sc = self.scalars.view(self.n_sys, self.stride)
        iters = sc[:, 7].to(torch.int64)
        rel = torch.sqrt(sc[:, 4] / torch.clamp(sc[:, 5], min=1e-300))
        self._last_iters = int(iters.max().item()) if self.n_sys else 0
        return self.x.clone(), iters, rel


_PCG_CACHE: dict = {}


def _solver_for(A, rtol: float) -> BatchedPCG:
    """One PCG workspace (and its captured graphs) per pattern; the solver
    keeps the pattern tensors alive, so their addresses identify it."""
    key = (A.row_ptr.data_ptr(), A.col_indices.data_ptr(), A.shape[0], A.values.dtype, rtol)
    s = _PCG_CACHE.get(key)
    if s is None:
        while len(_PCG_CACHE) >= 8:
            _PCG_CACHE.pop(next(iter(_PCG_CACHE)))
        s = BatchedPCG(A.row_ptr, A.col_indices, A.values, [0, A.shape[0]], rtol=rtol)
        s._keep = (A.row_ptr, A.col_indices)
        _PCG_CACHE[key] = s
    else:
        s.refresh_values(A.values)
    return s


def cg(A, b: torch.Tensor, x0: Optional[torch.Tensor] = None, rtol: float = 1e-10, atol: float = 0.0,
       max_iter: int = 100000, check_every: int = 25, refine: int = 0):
    """Jacobi-preconditioned conjugate gradients (torch_solve.py:54-128).

    Returns (x, iterations, relative residual).  Device matrices in CSR order
    use the native batched PCG with one system; others run in torch.

    ``refine`` (not in the reference): that many steps of iterative
    refinement after the solve, each solving A d = b - A x to ``rtol``
    relative to its own right-hand side and adding d.  The CG recurrence
    residual drifts from the true one, so a solve stopped at ``rtol`` leaves
    an error up to cond(A) * rtol; one refinement step takes it to the
    rounding level of the residual b - A x.  The reported residual is then
    the true |b - A x| / |b| and the iterations are the total.
    """
    x, it, rel = _cg(A, b, x0, rtol, atol, max_iter, check_every)
    _warn_unconverged(it, rel, rtol, atol, b)
    if refine <= 0:
        return x, it, rel
    b = b.detach().to(A.values.dtype)
    with torch.no_grad():
        for _ in range(refine):
            r = b - spmv(A, x)
            d, it_d, rel_d = _cg(A, r, None, rtol, 0.0, max_iter, check_every)
            _warn_unconverged(it_d, rel_d, rtol, 0.0, r)
            x = x + d
            it += it_d
        bnorm = torch.linalg.vector_norm(b).item()
        rel = torch.linalg.vector_norm(b - spmv(A, x)).item() / max(bnorm, 1e-300)
    return x, it, rel


def _cg(A, b, x0, rtol, atol, max_iter, check_every):
    if A.device.type == "cuda" and A.row_ptr is not None and b.dim() == 1:
        solver = _solver_for(A, rtol)
        x, it, rel = solver.solve(b.detach(), x0=None if x0 is None else x0.detach(), max_iter=max_iter)
        return x, int(it[0].item()), float(rel[0].item())
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


def _warn_unconverged(it: int, rel: float, rtol: float, atol: float, b: torch.Tensor) -> None:
    """Warn as the reference's CG does when the iteration budget runs out
    (torch_solve.py:125-126)."""
    bnorm = torch.linalg.vector_norm(b.detach()).item()
    if not (rel <= rtol or rel * bnorm <= atol):
        warnings.warn(f"cg did not converge after {it} iterations. with relative residual {rel:.3e}")


class _Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, b, A, kwargs):
        x, it, rel = cg(A, b, **kwargs)
        ctx.A, ctx.kwargs = A, kwargs
        ctx.save_for_backward(x)
        A.last_solve_info = (it, rel)
        return x

    @staticmethod
    def backward(ctx, gx):
        (x,) = ctx.saved_tensors
        A = ctx.A
        # symmetric A: A^T lam = gx is the forward system (torch_solve.py:262-272)
        lam, _, _ = cg(A, gx.contiguous(), **ctx.kwargs)
        gvals = -lam[A.row_indices] * x[A.col_indices]
        return gvals, lam, None, None


def solve(A, b: torch.Tensor, **kwargs) -> torch.Tensor:
    """x = A^-1 b for a symmetric positive definite A, differentiable in
    ``A.values`` and ``b``."""
    assert A.shape[0] == A.shape[1] == b.shape[0]
    kwargs.pop("verbose", None)
    return _Solve.apply(A.values, b, A, kwargs)
