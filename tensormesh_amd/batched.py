"""Many uneven P1 meshes assembled and solved together.

The reference has no batched assembly: each mesh goes through its own
``from_mesh`` / ``__call__`` / ``Condenser`` / ``solve`` in a Python loop
(docs/source/user_guide/batched_workflows.rst:33-35,253-268), with the
Jacobi-CG behaviour of tensormesh/sparse/solve/torch_solve.py:54-128.

Here the meshes are stacked into one disjoint union mesh: its CSR pattern is
block-diagonal, so one launch of the fused row kernel assembles every
stiffness matrix, one launch every load vector and one condensation pass
every Dirichlet reduction.  The condensed systems stay contiguous (mesh
k's inner dofs are a contiguous range), so the batched Jacobi-PCG iterates
all of them at once with per-system convergence flags, chunks of iterations
replayed from a HIP graph.

Precision (``mixed=True``): assembly and the CG bulk run in float32 (the
float32 kernels still evaluate element geometry and determinants in float64);
an outer iterative refinement forms the residual r = b - K x in float64 on
the device (float32 matrix values promoted exactly) and stops each system at
|r| <= rtol |b|.
"""

from __future__ import annotations

import time
from typing import List, Optional, Sequence

import numpy as np
import torch

from .builtin import DiffusionElementAssembler, SourceNodeAssembler
from .condense import Condenser
from .mesh import Mesh
from .solve import BatchedPCG


class BatchedDiffusion:
    """Stacked P1 variable-coefficient diffusion problems with homogeneous
    Dirichlet boundaries."""

    def __init__(self, meshes: Sequence[Mesh], quadrature_order: int = 2):
        assert len(meshes) > 0
        etype = meshes[0].default_element_type
        dev = meshes[0].device
        pts, cells, bnd = [], [], []
        self.node_off = [0]
        self.elem_off = [0]
        for m in meshes:
            assert m.default_element_type == etype and len(m.cells) == 1
            pts.append(m.points.to(torch.float64))
            cells.append(m.elements().to(dev) + self.node_off[-1])
            bnd.append(m.boundary_mask)
            self.node_off.append(self.node_off[-1] + m.n_points)
            self.elem_off.append(self.elem_off[-1] + m.elements().shape[0])
        self.union = Mesh(torch.cat(pts), {etype: torch.cat(cells)}, torch.cat(bnd))
        self.n_sys = len(meshes)
        self.K = DiffusionElementAssembler.from_mesh(self.union, quadrature_order=quadrature_order)
        self.K.p1_path = "auto"   # the P1 assembly path is timed once per union pattern (tensormesh_amd.tune)
        self.F = SourceNodeAssembler.from_mesh(self.union, quadrature_order=quadrature_order)
        self.cond = Condenser(self.union.boundary_mask)
        inner = (~self.union.boundary_mask).to(torch.int64).cpu()
        cs = torch.cumsum(inner, 0)
        self.sys_rows = [0] + [int(cs[o - 1]) for o in self.node_off[1:]]
        self._pcg = {}
        self._points32 = self.union.points.float()

    def assemble(self, kappa: torch.Tensor, f: torch.Tensor, dtype=torch.float64):
        """(K_ii, f_i) of every system, stacked: one assembly launch for all
        stiffness matrices, one for all loads, one condensation."""
        pts = self.union.points if dtype == torch.float64 else self._points32
        K = self.K(pts, element_data={"kappa": kappa.to(dtype)})
        b = self.F(pts, point_data={"f": f.to(dtype)})
        return self.cond(K, b)

    def _solver(self, Ki, rtol):
        key = (Ki.values.dtype, rtol)
        s = self._pcg.get(key)
        if s is None:
            s = BatchedPCG(Ki.row_ptr, Ki.col_indices, Ki.values, self.sys_rows, rtol=rtol)
            self._pcg[key] = s
        else:
            s.refresh_values(Ki.values)
        return s

    def solve(self, kappa: torch.Tensor, f: torch.Tensor, rtol: float = 1e-8, mixed: bool = True,
              inner_rtol: float = 1e-4, max_outer: int = 20):
        """Solve all systems; returns (u_full [sum N_k] float64, stats)."""
        t0 = time.perf_counter()
        if not mixed:
            Ki, fi = self.assemble(kappa, f, torch.float64)
            pcg = self._solver(Ki, rtol)
            x, it, rel = pcg.solve(fi)
            u = self.cond.recover(x)
            return u, {"outer": 1, "inner_iterations": it, "relres": rel, "seconds": time.perf_counter() - t0}
        Ki32, fi32 = self.assemble(kappa, f, torch.float32)
        # float64 copy of the float32-assembled operator: exact promotion, used for residuals
        Ki64 = Ki32.double()
        fi64 = self.F(self.union.points, point_data={"f": f.to(torch.float64)})[self.cond.is_inner_dof]
        res64 = self._solver(Ki64, rtol)
        inner = self._solver(Ki32, inner_rtol)
        x = torch.zeros_like(fi64)
        total_inner = torch.zeros(self.n_sys, dtype=torch.int64, device=x.device)
        outer = 0
        rel = None
        for outer in range(1, max_outer + 1):
            rel = res64.residual(fi64, x)                 # r = b - K x in float64, per-system |r|/|b|
            if bool((rel <= rtol).all().item()):
                outer -= 1
                break
            r = res64.r.clone()
            done = (rel <= rtol)
            if bool(done.any().item()):
                mask = torch.repeat_interleave(done, torch.tensor(np.diff(self.sys_rows), device=x.device))
                r = torch.where(mask, torch.zeros_like(r), r)
            d, it, _ = inner.solve(r.float())
            total_inner += it
            x = x + d.double()
        u = self.cond.recover(x)
        return u, {"outer": outer, "inner_iterations": total_inner, "relres": rel,
                   "seconds": time.perf_counter() - t0}
