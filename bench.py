#!/usr/bin/env python3
"""Benchmark of the tensorized element-assembly path on one MI355X.

Workload (one step): assemble the variable-coefficient P1 stiffness
K_ij = int kappa grad(phi_i).grad(phi_j) on the unit cube split into
64^3 cubes x 6 Kuhn tetrahedra = 1,572,864 tets / 274,625 nodes (interior
nodes jittered +-0.15h, seed 0), per-element kappa log-normal (mu=0, sigma=1,
seed 1), quadrature order 1, float64, through the public
``DiffusionElementAssembler(...)(element_data={'kappa': kappa})`` call; the
output is the CSR-ordered SparseMatrix (4,018,753 entries).  Inputs are
resident in HBM before the timed region; the pattern (from_mesh) is set up
once outside it and timed separately.

Prints ONE JSON line on stdout.  ``--gpus`` above 1 is refused: this build
targets one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12    # MI355X HBM3E, bytes/s
PEAK_FP64 = 78.6e12  # MI355X dense FP64 (vector and matrix), flop/s
PEAK_FP32 = 157.3e12  # MI355X dense FP32 matrix, flop/s


def algorithmic_bytes(E, N, nnz, B=4, D=3, s=8, c=8):
    """Unique bytes of the two-stage pipeline, int32 indices (every array read
    or written once):
       stage G = 4BE + cDN + sE + sB^2E      (conn, coords, kappa, slot values)
       stage S = 4B^2E + 4(nnz+1) + sB^2E + s*nnz   (seg_slot, seg_ptr, slots, values)"""
    g = 4 * B * E + c * D * N + s * E + s * B * B * E
    sred = 4 * B * B * E + 4 * (nnz + 1) + s * B * B * E + s * nnz
    return g, sred


def build_workload(dtype):
    import tensormesh_amd as tm
    from tensormesh_amd.mesh import Mesh, lognormal_coefficient, unit_cube_tet_arrays
    p, c, _ = unit_cube_tet_arrays(64, 0.15, 0)
    kap = lognormal_coefficient(c.shape[0], 1)
    mesh = Mesh(torch.from_numpy(p).to(dtype), {"tetra": torch.from_numpy(c)}).to("cuda")
    kappa = torch.from_numpy(kap).to(dtype=dtype, device="cuda")
    t0 = time.perf_counter()
    A = tm.DiffusionElementAssembler.from_mesh(mesh, quadrature_order=1)
    setup = time.perf_counter() - t0
    return mesh, kappa, A, setup, (p, c, kap)


def stage_times(A, mesh, kappa, reps=20):
    """Average device time of stage G and stage S (HIP events on the stream
    the kernels run on: the current torch stream)."""
    import ctypes
    from tensormesh_amd import _native as nat
    from tensormesh_amd.forms import FormMatch
    L = nat.lib()
    dev = mesh.points.device
    pts = mesh.points.contiguous()
    conn = A._conn32(dev)
    E, B = conn.shape
    pat = A._pattern_on(dev)
    tab, tab_dev = A._tables(dev)
    slot = torch.empty(E * B * B, dtype=pts.dtype, device=dev)
    vals = torch.empty(pat.nnz, dtype=pts.dtype, device=dev)
    mv = A._mesh_view(pts)
    pv = pat.view(pts.dtype, 1)
    fd = nat.FormDesc(nat.FORM["diffusion_elem"], 1, tab.ctypes.data, tab_dev.data_ptr(), kappa.data_ptr(), 0.0, 0.0)
    rv = pat.rows_view(conn)
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4 * reps)]
    for r in range(reps + 3):
        if r >= 3:
            ev[4 * (r - 3)].record(stream)
        nat.check(L.tm_element_matrices(ctypes.byref(mv), ctypes.byref(fd), slot.data_ptr(), s), "G")
        if r >= 3:
            ev[4 * (r - 3) + 1].record(stream)
        nat.check(L.tm_segmented_reduce(nat.dtype_code(pts.dtype), pat.nnz, 1, pat.seg_ptr.data_ptr(),
                                        pat.seg_slot.data_ptr(), pv.n_blocks, pv.blk_ptr, slot.data_ptr(),
                                        vals.data_ptr(), s), "S")
        if r >= 3:
            ev[4 * (r - 3) + 2].record(stream)
        nat.check(L.tm_reduce_p1_rows(nat.dtype_code(pts.dtype), nat.CELL["tetra"], ctypes.byref(rv),
                                      slot.data_ptr(), 0, vals.data_ptr(), s), "S_rows")
        if r >= 3:
            ev[4 * (r - 3) + 3].record(stream)
    torch.cuda.synchronize()
    g = np.mean([ev[4 * r].elapsed_time(ev[4 * r + 1]) for r in range(reps)]) * 1e-3
    sr = np.mean([ev[4 * r + 1].elapsed_time(ev[4 * r + 2]) for r in range(reps)]) * 1e-3
    sr2 = np.mean([ev[4 * r + 2].elapsed_time(ev[4 * r + 3]) for r in range(reps)]) * 1e-3
    return g, sr, sr2


def fused_time(A, mesh, kappa, reps=20):
    """Average device time of the fused row-centric kernel (HIP events on the
    current stream)."""
    import ctypes
    from tensormesh_amd import _native as nat
    L = nat.lib()
    dev = mesh.points.device
    pts = mesh.points.contiguous()
    conn = A._conn32(dev)
    pat = A._pattern_on(dev)
    tab, tab_dev = A._tables(dev)
    vals = torch.empty(pat.nnz, dtype=pts.dtype, device=dev)
    mv = A._mesh_view(pts)
    rv = pat.rows_view(conn)
    fd = nat.FormDesc(nat.FORM["diffusion_elem"], 1, tab.ctypes.data, tab_dev.data_ptr(), kappa.data_ptr(), 0.0, 0.0)
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    for r in range(3):
        nat.check(L.tm_assemble_p1_rows(ctypes.byref(mv), ctypes.byref(fd), ctypes.byref(rv), vals.data_ptr(),
                                        stream.cuda_stream), "rows")
    ev[0].record(stream)
    for r in range(reps):
        nat.check(L.tm_assemble_p1_rows(ctypes.byref(mv), ctypes.byref(fd), ctypes.byref(rv), vals.data_ptr(),
                                        stream.cuda_stream), "rows")
        ev[r + 1].record(stream)
    torch.cuda.synchronize()
    return float(np.mean([ev[r].elapsed_time(ev[r + 1]) for r in range(reps)]) * 1e-3), rv.n_blocks


def incidence_stage_times(A, mesh, kappa, reps=20):
    """Device time of the incidence-ordered path's two kernels (HIP events on
    the current stream): stage G writing rows in incidence order, stage S
    reducing contiguous rows."""
    import ctypes
    from tensormesh_amd import _native as nat
    L = nat.lib()
    dev = mesh.points.device
    pts = mesh.points.contiguous()
    conn = A._conn32(dev)
    E, B = conn.shape
    pat = A._pattern_on(dev)
    tab, tab_dev = A._tables(dev)
    slot = torch.empty(E * B * B, dtype=pts.dtype, device=dev)
    vals = torch.empty(pat.nnz, dtype=pts.dtype, device=dev)
    mv = A._mesh_view(pts)
    rv = pat.rows_view(conn)
    inv = pat.inc_inv()
    fd = nat.FormDesc(nat.FORM["diffusion_elem"], 1, tab.ctypes.data, tab_dev.data_ptr(), kappa.data_ptr(), 0.0, 0.0)
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3 * reps)]
    for r in range(reps + 3):
        if r >= 3:
            ev[3 * (r - 3)].record(stream)
        nat.check(L.tm_element_matrices_inc(ctypes.byref(mv), ctypes.byref(fd), inv.data_ptr(), slot.data_ptr(), s),
                  "G_inc")
        if r >= 3:
            ev[3 * (r - 3) + 1].record(stream)
        nat.check(L.tm_reduce_p1_rows(nat.dtype_code(pts.dtype), nat.CELL["tetra"], ctypes.byref(rv),
                                      slot.data_ptr(), 1, vals.data_ptr(), s), "S_inc")
        if r >= 3:
            ev[3 * (r - 3) + 2].record(stream)
    torch.cuda.synchronize()
    g = float(np.mean([ev[3 * r].elapsed_time(ev[3 * r + 1]) for r in range(reps)]) * 1e-3)
    sr = float(np.mean([ev[3 * r + 1].elapsed_time(ev[3 * r + 2]) for r in range(reps)]) * 1e-3)
    return g, sr


def fused_bytes(E, N, nnz, B=4, D=3, s=8, c=8):
    """Unique bytes of the fused row-centric kernel, int32 indices:
    conn 4BE + coords cDN + kappa sE + incidence 4BE + colpos 4BE + row_ptr and
    inc_ptr 8(N+1) + entry masks 4 nnz + values s*nnz."""
    return 4 * B * E + c * D * N + s * E + 4 * B * E + 4 * B * E + 8 * (N + 1) + 4 * nnz + s * nnz


def rows_reduce_bytes(E, N, nnz, B=4, s=8):
    """Unique bytes of the row-centric reduction: incidence 4BE + colpos 4BE +
    slot values sB^2E + row_ptr / inc_ptr 8(N+1) + entry masks 4 nnz + values s*nnz."""
    return 4 * B * E + 4 * B * E + s * B * B * E + 8 * (N + 1) + 4 * nnz + s * nnz


def cpu_baseline(p, c, kap, budget_s=25.0):
    """The oracle (torch-CPU float64 restatement of the reference chain:
    gather, einsum Jacobian, batched inverse / det, einsum form, integrate,
    index_add_) on the full workload, host cores, median of a few runs."""
    from oracle import fem
    # the host CPUs this process may run on; a GPU job's share of the host
    # (OMP_NUM_THREADS where the launcher sets it) caps the thread count
    avail = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", avail) or avail)
    ncores = max(1, min(avail, share))
    torch.set_num_threads(ncores)
    P, C, K = torch.from_numpy(p), torch.from_numpy(c), torch.from_numpy(kap)
    _, _, eids = fem.pattern(c, p.shape[0])
    eids = torch.from_numpy(eids)
    nnz = int(eids.max()) + 1

    def run():
        loc = fem.element_integrals(P, C, "tetra", 1, "diffusion_elem", kappa=K)
        return fem._index_add(loc, eids, nnz)

    t0 = time.perf_counter()
    run()
    first = time.perf_counter() - t0
    times = []
    n = max(2, min(7, int(budget_s / max(first, 1e-3))))
    for _ in range(n):
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {"value": c.shape[0] / med, "unit": "elements/s", "cores": ncores, "kind": "port",
            "sample": f"full 1,572,864-tet workload, oracle op sequence, median of {n} runs after 1 warm-up "
                      f"({med:.3f} s per assembly, torch CPU threads={ncores} of {avail} CPUs in the "
                      f"process affinity, {os.cpu_count()} on the host)",
            "seconds_per_assembly": med, "cpus_in_affinity": avail, "host_cpus": os.cpu_count()}


def _reference_pcg(row, col, values, n, b, rtol, max_iter=100000):
    """Jacobi-preconditioned conjugate gradients in torch ops on a torch CSR
    matrix: the method SparseMatrix.solve runs on the device through the
    torch-sla package with its defaults (tensormesh/sparse/solve/__init__.py:42-
    45, 127-160: backend 'pytorch', method 'cg', preconditioner 'jacobi'; the
    package is not installed here, so this restates the method), written as
    the reference's own torch solvers are (torch_solve.py:90-136): one torch
    op per vector step and a convergence test on |r| every iteration.  Stops
    at |r| <= rtol |b|.  (The solver the reference falls back to without
    torch-sla, SparseSolveTorch's BiCGSTAB, stops on its absolute breakdown
    test after 51 iterations on this system, far from converged.)"""
    A = torch.sparse_coo_tensor(torch.stack([row, col]), values, (n, n)).coalesce().to_sparse_csr()
    diag = torch.ones(n, dtype=values.dtype, device=values.device)
    on = row == col
    diag[row[on]] = values[on]
    minv = 1.0 / diag
    x = torch.zeros_like(b)
    r = b.clone()
    z = minv * r
    p = z.clone()
    rz = torch.dot(r, z)
    tol = rtol * torch.norm(b)
    it = 0
    for it in range(1, max_iter + 1):
        Ap = A @ p
        alpha = rz / torch.dot(p, Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        if torch.norm(r) <= tol:
            break
        z = minv * r
        rz_new = torch.dot(r, z)
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, it


class _ReferenceSolve(torch.autograd.Function):
    """Differentiable solve of the reference path: forward solve, backward the
    transposed solve and values grad -lambda[row] * u[col] (the adjoint of
    torch_solve.py:262-272)."""

    @staticmethod
    def forward(ctx, values, row, col, n, b, rtol):
        u, it = _reference_pcg(row, col, values.detach(), n, b.detach(), rtol)
        ctx.save_for_backward(values, row, col, u)
        ctx.n, ctx.rtol = n, rtol
        _ReferenceSolve.last_iterations = it
        return u

    @staticmethod
    def backward(ctx, g):
        values, row, col, u = ctx.saved_tensors
        lam, _ = _reference_pcg(col, row, values.detach(), ctx.n, g, ctx.rtol)
        return -lam[row] * u[col], None, None, None, lam, None


def grad_step_bench(n=512, steps=2, rtol=1e-10):
    """Forward + adjoint gradient step: 2-D P1 diffusion on an n x n x 2
    triangle mesh (jitter +-0.2h, seed 0), kappa log-normal per element
    (seed 1), loss ||K(kappa, x)^-1 b - u_target||^2 with u_target from a
    kappa field of seed 2, homogeneous Dirichlet condensation, b the load of
    f = 2 pi^2 sin(pi x) sin(pi y) (held fixed).  dL/dkappa and dL/dx.

    ours : HIP assembly + hand-written adjoint kernel, native Jacobi-PCG
           (two launches per iteration) with its adjoint solve
    plain: the restated reference path on the same GPU -- vmap'd torch form
           + autograd through einsum / inverse / det / index_add, and a
           torch-op Jacobi-PCG with its adjoint backward (the method the
           reference's SparseMatrix.solve runs on the device)
    Both stop at the same residual, |r| <= rtol |b|, and use the same
    condensation.  "plain_native_solve" is the reference assembly with the
    native solver, which isolates the assembly share."""
    import tensormesh_amd as tm
    from tensormesh_amd.mesh import lognormal_coefficient, unit_square_tri
    m = unit_square_tri(n, 0.2, 0).to("cuda")
    E = m.elements().shape[0]
    kap = torch.from_numpy(lognormal_coefficient(E, 1)).cuda()
    kap_t = torch.from_numpy(lognormal_coefficient(E, 2)).cuda()
    A = tm.DiffusionElementAssembler.from_mesh(m)
    P = tm.DiffusionElementAssembler.from_assembler(A)
    P.reference_path = True
    cond = tm.Condenser(m.boundary_mask)
    x, y = m.points[:, 0], m.points[:, 1]
    f = 2 * np.pi ** 2 * torch.sin(np.pi * x) * torch.sin(np.pi * y)
    b = tm.SourceNodeAssembler.from_mesh(m)(point_data={"f": f})
    with torch.no_grad():
        Kt = A(element_data={"kappa": kap_t})
        Kti, fti = cond(Kt, b)
        u_t = Kti.solve(fti, rtol=rtol)

    def step(asm, reference_solver=False):
        sync = torch.cuda.synchronize
        kr = kap.clone().requires_grad_(True)
        pr = m.points.clone().requires_grad_(True)
        sync(); t0 = time.perf_counter()
        K = asm(pr, element_data={"kappa": kr})
        Ki, fi = cond(K, b)
        sync(); t1 = time.perf_counter()
        if reference_solver:
            u = _ReferenceSolve.apply(Ki.values, Ki.row, Ki.col, Ki.shape[0], fi, rtol)
            it_fwd = _ReferenceSolve.last_iterations
        else:
            u = Ki.solve(fi, rtol=rtol)
            it_fwd = Ki.last_solve_info[0]
        loss = ((u - u_t) ** 2).sum()
        sync(); t2 = time.perf_counter()
        loss.backward()
        sync(); t3 = time.perf_counter()
        return {"total_ms": (t3 - t0) * 1e3, "assembly_ms": (t1 - t0) * 1e3, "solve_ms": (t2 - t1) * 1e3,
                "backward_ms": (t3 - t2) * 1e3, "cg_iterations": it_fwd, "loss": loss.item(),
                "grads": (kr.grad.detach(), pr.grad.detach())}

    res = {}
    for name, asm, ref_solver, n_steps in (("ours", A, False, steps), ("plain", P, True, 1),
                                           ("plain_native_solve", P, False, steps)):
        step(asm, ref_solver)   # warm-up (pattern plans, graph capture, allocator)
        runs = [step(asm, ref_solver) for _ in range(n_steps)]
        res[name] = min(runs, key=lambda r: r["total_ms"])
    gk_o, gx_o = res["ours"].pop("grads")
    gk_p, gx_p = res["plain"].pop("grads")
    res["plain_native_solve"].pop("grads")
    agree = max(((gk_o - gk_p).abs().max() / gk_p.abs().max()).item(),
                ((gx_o - gx_p).abs().max() / gx_p.abs().max()).item())
    out = {"workload": f"P1 tri diffusion {n}x{n}x2 = {E} triangles, {m.n_points} nodes, dL/dkappa and dL/dx, "
                       f"solves to |r| <= {rtol:g} |b|",
           "ours": res["ours"], "plain": res["plain"], "plain_native_solve": res["plain_native_solve"],
           # primary: the reference assembly + autograd with the same solver as ours
           "speedup_same_solver": res["plain_native_solve"]["total_ms"] / res["ours"]["total_ms"],
           # secondary: against the restated torch-op Jacobi-PCG (torch-sla's device default is
           # not installed here; that restatement is parity unpinned)
           "speedup_vs_restated_torch_solver": res["plain"]["total_ms"] / res["ours"]["total_ms"],
           "solve_share_ours": (res["ours"]["solve_ms"] + res["ours"]["backward_ms"]) / res["ours"]["total_ms"],
           "gradient_agreement_max_rel": agree}
    with torch.no_grad():
        Ki, fi = cond(A(element_data={"kappa": kap}), b)
    out["pcg_us_per_iteration"] = pcg_iteration_bench(Ki, fi)
    return out


def pcg_iteration_bench(Ki, fi, iters=512, chunk=64):
    """Device time per Jacobi-PCG iteration on one condensed system, per
    variant (launches per iteration 1 / 2 / 5), from HIP events around graph
    replays of ``chunk`` iterations with a tolerance no iterate reaches."""
    from tensormesh_amd.solve import BatchedPCG
    out = {"rows": int(Ki.shape[0]), "nnz": int(Ki.nnz)}
    saved = BatchedPCG.launches
    try:
        for launches in (2, 1, 5):
            BatchedPCG.launches = launches
            s = BatchedPCG(Ki.row_ptr, Ki.col_indices, Ki.values, [0, Ki.shape[0]], rtol=1e-300)
            s.solve(fi, max_iter=chunk, chunk=chunk)      # captures the graph
            torch.cuda.synchronize()
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            s.residual(fi, torch.zeros_like(fi))   # restart from x = 0, parity 0
            g = s._graphs[chunk]
            ev[0].record()
            for _ in range(iters // chunk):
                g.replay()
            ev[1].record()
            torch.cuda.synchronize()
            out[f"launches_{launches}"] = ev[0].elapsed_time(ev[1]) * 1e3 / iters
    finally:
        BatchedPCG.launches = saved
    return out


def hyperelastic_bench(n=32, reps=20):
    """Neo-Hookean residual + consistent tangent per Newton step on the unit
    cube split into n^3 x 6 Kuhn tets promoted to tetra10 (32^3 -> 196,608
    elements, 274,625 nodes, 823,875 dofs), E = 1, nu = 0.3, 4-point rule,
    at the displacement u = 0.1 x (the prescribed face value spread linearly).
    Times float64 and float32 assembly (device time via HIP events): the
    default slot path, the same with symmetric and with full element blocks
    forced (``*_symmetric_blocks``, ``*_full_blocks``), the block-row path
    (``*_rows``) and the float32 MFMA tangent."""
    import tensormesh_amd as tm
    from tensormesh_amd.mesh import unit_cube_tet
    m = unit_cube_tet(n, 0.0, 0, order=2).to("cuda")
    out = {"workload": f"Neo-Hookean tetra10 residual + 30x30 tangent, {n}^3x6 = {m.elements().shape[0]} elements, "
                       f"{m.n_points} nodes, {3 * m.n_points} dofs, qorder 2"}
    t0 = time.perf_counter()
    model = tm.NeoHookeanModel.from_mesh(m, quadrature_order=2, E=1.0, nu=0.3)
    out["pattern_setup_s"] = time.perf_counter() - t0
    _warm(lambda: model.residual_and_tangent(torch.zeros_like(m.points)))
    for key, dt, mfma, path, sym in (("f64", torch.float64, False, "slots", None),
                                     ("f64_symmetric_blocks", torch.float64, False, "slots", True),
                                     ("f64_full_blocks", torch.float64, False, "slots", False),
                                     ("f64_rows", torch.float64, False, "rows", None),
                                     ("f32", torch.float32, False, "slots", None),
                                     ("f32_symmetric_blocks", torch.float32, False, "slots", True),
                                     ("f32_full_blocks", torch.float32, False, "slots", False),
                                     ("f32_rows", torch.float32, False, "rows", None),
                                     ("f32_mfma", torch.float32, True, "slots", None)):
        model.use_mfma = mfma
        model.nh_path = path
        model.symmetric_slots = sym
        pts = m.points.to(dt)
        u = torch.zeros_like(pts)
        u[:, 0] = 0.1 * pts[:, 0]
        # every variant warmed alike (clocks, caches, plans), untimed
        _warm(lambda: model.residual_and_tangent(u, points=pts))
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(reps):
            R, K = model.residual_and_tangent(u, points=pts)
        ev[1].record()
        torch.cuda.synchronize()
        ms = ev[0].elapsed_time(ev[1]) / reps
        E = m.elements().shape[0]
        s = 8 if dt == torch.float64 else 4
        nnzb = model.pattern.nnz
        # element kernel writes E*900 + E*30 values; the block reduction reads them plus indices
        byts = (4 * 10 * E + s * 3 * 2 * m.n_points + s * E * 930) + (4 * 100 * E + 4 * (nnzb + 1) + s * E * 900
                                                                       + s * nnzb * 9) \
            + (4 * 10 * E + 4 * (m.n_points + 1) + s * E * 30 + s * 3 * m.n_points)
        # tangent flops of the dense contraction B^T A B in F-space, 4 points x (2*9*9*30 + 2*30*30*9)
        flops = 84240 * E
        peak = PEAK_FP64 if dt == torch.float64 else PEAK_FP32
        out[key] = {
            "ms_per_newton_assembly": ms, "elements_per_s": E / (ms * 1e-3), "block_entries": nnzb,
            "scalar_nnz": nnzb * 9, "algorithmic_bytes": byts, "GBps": byts / (ms * 1e-3) / 1e9,
            "hbm_frac": byts / (ms * 1e-3) / PEAK_HBM,
            "tangent_flops": flops, "TFLOPs": flops / (ms * 1e-3) / 1e12,
            "flop_frac": flops / (ms * 1e-3) / peak}
    out["energy_backward"] = neohookean_energy_bench(m, model, reps)
    return out


def neohookean_energy_bench(m, model, reps=20):
    """Neo-Hookean energy(u).backward() at u = (0.1 x, 0.01 sin(pi x) sin(pi z), 0),
    float64: the HIP energy
    with its residual backward against the reference algorithm (vmapped torch
    energy + autograd) on the same GPU; device time via HIP events."""
    pts = m.points
    u0 = torch.zeros_like(pts)
    u0[:, 0] = 0.1 * pts[:, 0]
    # a non-uniform part, so that elements differ
    u0[:, 1] = 0.01 * torch.sin(np.pi * pts[:, 0]) * torch.sin(np.pi * pts[:, 2])

    def run(reference):
        model.reference_path = reference
        u = u0.clone().requires_grad_(True)
        e = model.energy(u)
        e.backward()
        return e.detach(), u.grad

    res = {}
    for name, ref, n in (("ours", False, reps), ("reference_algorithm", True, 3)):
        e, g = run(ref)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(n):
            e, g = run(ref)
        ev[1].record()
        torch.cuda.synchronize()
        res[name] = {"ms": ev[0].elapsed_time(ev[1]) / n, "energy": e.item(), "grad": g}
    model.reference_path = False
    gd = (res["ours"].pop("grad") - res["reference_algorithm"]["grad"]).abs().max().item()
    gs = res["reference_algorithm"].pop("grad").abs().max().item()
    return {"ours_ms": res["ours"]["ms"], "reference_algorithm_ms": res["reference_algorithm"]["ms"],
            "speedup": res["reference_algorithm"]["ms"] / res["ours"]["ms"],
            "energy_rel_diff": abs(res["ours"]["energy"] - res["reference_algorithm"]["energy"])
            / abs(res["reference_algorithm"]["energy"]),
            "gradient_max_rel_diff": gd / gs}


def facet_bench(n=64, reps=20):
    """Boundary-facet load (FacetAssembler, f * v with nodal f) on the
    boundary of the 64^3 x 6 Kuhn tetrahedral cube, float64, qorder 2: the HIP
    kernel against the reference algorithm (the vmapped torch path, as for any
    other integrand) on the same GPU, device time per call via HIP events."""
    import tensormesh_amd as tm
    from tensormesh_amd.mesh import Mesh, unit_cube_tet_arrays
    p, c, bm = unit_cube_tet_arrays(n, 0.15, 0)
    P, C, BM = torch.from_numpy(p), torch.from_numpy(c), torch.from_numpy(bm)
    m = Mesh(P, {"tetra": C}, BM).to("cuda")

    class Load(tm.FacetAssembler):
        def forward(self, v, f):
            return f * v

    A = Load.from_mesh(m, quadrature_order=2)
    f = (torch.sin(3 * P[:, 0]) + P[:, 1]).to("cuda")
    pd = {"f": f, "x": m.points}

    def timed(call):
        out = call()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(reps):
            out = call()
        ev[1].record()
        torch.cuda.synchronize()
        return out, ev[0].elapsed_time(ev[1]) / reps

    b, ms = timed(lambda: A(point_data={"f": f}))
    b_ref, ms_ref = timed(lambda: A._torch(A.forward, m.points, pd))
    F = int(A._topo.elem.numel())
    return {"workload": f"facet load f*v on the boundary of {n}^3x6 tets: {F} facets "
                        f"(incl. interior faces with all nodes on the boundary, as the reference selects)",
            "ms_per_call": ms, "facets_per_s": F / (ms * 1e-3), "reference_algorithm_ms_per_call": ms_ref,
            "speedup": ms_ref / ms, "max_rel_diff": float((b - b_ref).abs().max() / b_ref.abs().max())}


def plasticity_bench(n=48, reps=10):
    """J2 plasticity (J2Plasticity, E = 200, nu = 0.3, sig0 = 0.25, H = 1) on
    the jittered n^3 x 6 Kuhn tetrahedral cube (48^3 -> 663,552 tets), P1,
    4-point rule, float64, at a displacement past yield at most points with a
    nonzero prior plastic state.  Device time per call (HIP events) of the
    native energy, energy + backward (the reference's residual workflow),
    residual + consistent tangent and update_state, beside the reference
    algorithm (generic vmapped torch energy + autograd backward) on the same
    GPU."""
    import tensormesh_amd as tm
    from tensormesh_amd.mesh import unit_cube_tet
    m = unit_cube_tet(n, 0.1, 0).to("cuda")
    model = tm.J2Plasticity.from_mesh(m, quadrature_order=2, E=200.0, nu=0.3, sig0=0.25, H=1.0)
    P = m.points
    u = torch.stack([3e-3 * P[:, 0] + 2e-4 * torch.sin(6 * P[:, 1]), -1e-3 * P[:, 1],
                     5e-4 * torch.cos(5 * P[:, 0]) * P[:, 2]], 1)
    g = torch.Generator(device="cuda").manual_seed(0)
    h = model.history["tetra"]
    h["eps_p"].copy_(1e-4 * torch.randn(h["eps_p"].shape, generator=g, device="cuda", dtype=torch.float64))
    h["eps_p"].copy_(0.5 * (h["eps_p"] + h["eps_p"].transpose(-1, -2)))
    h["alpha"].copy_(1e-4 * torch.rand(h["alpha"].shape, generator=g, device="cuda", dtype=torch.float64))
    hist = {"eps_p_n": h["eps_p"], "alpha_n": h["alpha"]}

    def timed(call, n_rep=reps):
        out = call()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(n_rep):
            out = call()
        ev[1].record()
        torch.cuda.synchronize()
        return out, ev[0].elapsed_time(ev[1]) / n_rep

    def energy_and_grad():
        x = u.clone().requires_grad_(True)
        e = model.energy(point_data={"displacement": x}, element_data=hist)
        e.backward()
        return e.detach(), x.grad

    _warm(lambda: model.residual_and_tangent(u))
    e_nat, ms_e = timed(lambda: model.energy(point_data={"displacement": u}, element_data=hist))
    (e1, g_nat), ms_eg = timed(energy_and_grad)
    _, ms_rt = timed(lambda: model.residual_and_tangent(u))
    alt = {}
    for name, sym in (("symmetric_blocks", True), ("full_blocks", False)):
        model.symmetric_slots = sym
        alt[f"ms_residual_and_tangent_{name}"] = timed(lambda: model.residual_and_tangent(u))[1]
    model.symmetric_slots = None
    model.reference_path = True
    (e_ref, g_ref), ms_ref = timed(energy_and_grad, 3)
    model.reference_path = False
    saved = (h["eps_p"].clone(), h["alpha"].clone())
    _, ms_up = timed(lambda: model.update_state(u), 1)
    h["eps_p"].copy_(saved[0])
    h["alpha"].copy_(saved[1])
    E = m.elements().shape[0]
    return {"workload": f"J2 plasticity P1 tets {n}^3x6 = {E} elements, {m.n_points} nodes, qorder 2 (4 points), f64",
            "ms_energy": ms_e, "ms_energy_and_backward": ms_eg, "ms_residual_and_tangent": ms_rt,
            **alt, "ms_update_state": ms_up, "reference_algorithm_ms_energy_and_backward": ms_ref,
            "speedup_energy_and_backward": ms_ref / ms_eg,
            "elements_per_s_residual_and_tangent": E / (ms_rt * 1e-3),
            "energy_rel_diff": abs(e_nat.item() - e_ref.item()) / abs(e_ref.item()),
            "grad_max_rel_diff": float((g_nat - g_ref).abs().max() / g_ref.abs().max())}


def _warm(call, seconds=0.2):
    """Repeat ``call`` untimed for about ``seconds`` so that the first timed
    variant of a leg does not pay the device's clock ramp."""
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        call()
        torch.cuda.synchronize()


def _timed(call, reps):
    out = call()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(reps):
        out = call()
    ev[1].record()
    torch.cuda.synchronize()
    return out, ev[0].elapsed_time(ev[1]) / reps


def contact_bench(n=64, reps=20):
    """Penalty contact (PenaltyContact, kappa = 1e3, plane y = 0.02) on the
    boundary of the jittered 64^3 x 6 Kuhn tetrahedral cube pushed 0.05 below
    the plane, float64, qorder 2: energy + backward (the reference's workflow)
    natively and by the reference algorithm (vmapped torch density +
    autograd) on the same GPU, and the native gradient + Hessian."""
    import tensormesh_amd as tm
    from tensormesh_amd.mesh import Mesh, unit_cube_tet_arrays
    p, c, bm = unit_cube_tet_arrays(n, 0.15, 0)
    m = Mesh(torch.from_numpy(p), {"tetra": torch.from_numpy(c)}, torch.from_numpy(bm)).to("cuda")
    model = tm.PenaltyContact.from_mesh(m, quadrature_order=2, kappa=1e3, obstacle_y=0.02)
    x = m.points.clone()
    x[:, 1] -= 0.05

    def energy_and_grad():
        xx = x.clone().requires_grad_(True)
        e = model.energy(points=xx)
        e.backward()
        return e.detach(), xx.grad

    (e_nat, g_nat), ms = _timed(energy_and_grad, reps)
    _, ms_rt = _timed(lambda: model.residual_and_tangent(x), reps)
    model.reference_path = True
    (e_ref, g_ref), ms_ref = _timed(energy_and_grad, 5)
    model.reference_path = False
    F = int(model._topo.elem.numel())
    return {"workload": f"penalty contact on the boundary of {n}^3x6 tets: {F} facets, f64, qorder 2",
            "ms_energy_and_backward": ms, "ms_gradient_and_hessian": ms_rt,
            "reference_algorithm_ms_energy_and_backward": ms_ref, "speedup_energy_and_backward": ms_ref / ms,
            "energy_rel_diff": abs(e_nat.item() - e_ref.item()) / abs(e_ref.item()),
            "grad_max_rel_diff": float((g_nat - g_ref).abs().max() / g_ref.abs().max())}


def hex_bench(n=64, reps=10):
    """Hexahedral assembly on the jittered 64^3 hexahedron grid (262,144 cells,
    274,625 nodes), float64, 2 x 2 x 2 points: Laplace stiffness and linear
    elasticity (3 x 3 blocks) natively and by the reference algorithm (the
    vmapped torch integrand and index_add projection) on the same GPU."""
    import tensormesh_amd as tm
    from tensormesh_amd.mesh import unit_cube_hex
    m = unit_cube_hex(n, 0.15, 0).to("cuda")
    out = {"workload": f"hexahedron {n}^3 = {m.elements().shape[0]} cells, {m.n_points} nodes, qorder 2 (8 points), f64"}
    for name, cls, kw in (("laplace", tm.LaplaceElementAssembler, {}),
                          ("elasticity", tm.LinearElasticityElementAssembler, {"E": 1.0, "nu": 0.3})):
        A = cls.from_mesh(m, quadrature_order=2, **kw)
        _warm(lambda: A())
        K, ms = _timed(lambda: A(), reps)
        alt = {}
        for alt_name, sym in (("symmetric_blocks", True), ("full_blocks", False)):
            A.symmetric_slots = sym
            alt[f"ms_per_call_{alt_name}"] = _timed(lambda: A(), reps)[1]
        A.symmetric_slots = None
        A.reference_path = True
        K_ref, ms_ref = _timed(lambda: A(), 2)
        A.reference_path = False
        out[name] = {"ms_per_call": ms, **alt,
                     "cells_per_s": m.elements().shape[0] / (ms * 1e-3),
                     "reference_algorithm_ms_per_call": ms_ref, "speedup": ms_ref / ms,
                     "max_rel_diff": float((K.values - K_ref.values).abs().max() / K_ref.values.abs().max())}
    return out


def batched_bench(n_meshes=256, seed=0):
    """256 jittered 2-D P1 meshes, side N ~ U{16..128} (seeded), own log-normal
    kappa and a sinusoidal source each; one assembly launch for all stiffness
    matrices and loads, batched Jacobi-PCG to 1e-8 relative residual, float32
    bulk with float64 iterative refinement, iteration chunks replayed from a
    HIP graph.  Compared with solving the same problems one by one through the
    same API (float64 PCG per mesh)."""
    from tensormesh_amd.batched import BatchedDiffusion
    from tensormesh_amd.mesh import lognormal_coefficient, unit_square_tri
    rng = np.random.default_rng(seed)
    sizes = rng.integers(16, 129, size=n_meshes)
    meshes = [unit_square_tri(int(nk), 0.2, 1000 + k).to("cuda") for k, nk in enumerate(sizes)]
    kap = torch.cat([torch.from_numpy(lognormal_coefficient(mk.elements().shape[0], 2000 + k))
                     for k, mk in enumerate(meshes)]).cuda()
    f = torch.cat([torch.sin(np.pi * mk.points[:, 0]) * torch.sin(np.pi * mk.points[:, 1]) * (1 + k % 3)
                   for k, mk in enumerate(meshes)])
    t0 = time.perf_counter()
    bp = BatchedDiffusion(meshes)
    setup = time.perf_counter() - t0
    bp.solve(kap, f, rtol=1e-8, mixed=True)      # warm-up: plans, graph capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u, st = bp.solve(kap, f, rtol=1e-8, mixed=True)
    torch.cuda.synchronize()
    t_b = time.perf_counter() - t0
    u64, st64 = bp.solve(kap, f, rtol=1e-10, mixed=False)
    err = ((u - u64).norm() / u64.norm()).item()
    # one by one through the same API (a per-mesh loop as in the reference's workflow)
    import tensormesh_amd as tm
    singles = []
    for k, mk in enumerate(meshes[:32]):
        A = tm.DiffusionElementAssembler.from_mesh(mk)
        F = tm.SourceNodeAssembler.from_mesh(mk)
        cd = tm.Condenser(mk.boundary_mask)
        singles.append((A, F, cd, kap[bp.elem_off[k]:bp.elem_off[k + 1]], f[bp.node_off[k]:bp.node_off[k + 1]]))

    def one(A, F, cd, kk, ff):
        K = A(element_data={"kappa": kk})
        b = F(point_data={"f": ff})
        Ki, fi = cd(K, b)
        return Ki.solve(fi, rtol=1e-8)

    for item in singles:
        one(*item)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for item in singles:
        one(*item)
    torch.cuda.synchronize()
    t_loop = (time.perf_counter() - t0) / len(singles) * n_meshes
    return {"workload": f"{n_meshes} P1 meshes, side ~U{{16..128}}, {bp.elem_off[-1]} triangles, "
                        f"{bp.node_off[-1]} nodes in total; rtol 1e-8",
            "batched_s": t_b, "problems_per_s": n_meshes / t_b, "setup_s": setup,
            "outer_refinements": st["outer"], "max_inner_iterations": int(st["inner_iterations"].max()),
            "max_relres": float(st["relres"].max()), "mixed_vs_fp64_solution_rel": err,
            "loop_same_api_s_extrapolated": t_loop, "speedup_vs_loop": t_loop / t_b}


# the sources that decide what the headline kernel (p1_rows_dest_kernel) reads and
# writes: the kernel, its headers and the builders of the plans it walks
ROOFLINE_SOURCES = ("tensormesh_amd/csrc/tm_rows.hip", "tensormesh_amd/csrc/tm_device.h",
                    "tensormesh_amd/csrc/tm_internal.h", "include/tensormesh_amd.h",
                    "tensormesh_amd/csrc/tm_pattern.cpp", "tensormesh_amd/csrc/tm_pattern_gpu.hip",
                    "tensormesh_amd/pattern.py")


def source_digest(paths=ROOFLINE_SOURCES):
    """SHA-256 of the given sources: ties counter data to the code it was
    measured on."""
    import hashlib
    h = hashlib.sha256()
    for rel in paths:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(rel.encode() + f.read())
    return h.hexdigest()


def fp32_bench(steps=30, warmup=5):
    """The headline workload in float32 (coordinates, kappa and values float32;
    the fused kernel evaluates element geometry in float32, redoing slivers in
    float64, and accumulates in float32): elements/s, the fraction of
    the HBM roofline on the float32 two-stage pipeline bytes, and the error
    against the float64 result of the same call on the same (float32)
    inputs, which the GPU tests hold to the float64 oracle within 1e-12."""
    mesh32, kap32, A32, _, _ = build_workload(torch.float32)
    E = A32.elements["tetra"].shape[0]
    N, nnz = mesh32.n_points, A32.pattern.nnz
    ed = {"kappa": kap32}
    for _ in range(warmup):
        K32 = A32(element_data=ed)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        K32 = A32(element_data=ed)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    k_s, _ = fused_time(A32, mesh32, kap32)
    import tensormesh_amd as tm
    from tensormesh_amd.mesh import Mesh
    m64 = Mesh(mesh32.points.double(), mesh32.cells)
    K64 = tm.DiffusionElementAssembler.from_assembler(A32)(m64.points, element_data={"kappa": kap32.double()})
    d = K32.values.double() - K64.values
    gb, sb = algorithmic_bytes(E, N, nnz, s=4, c=4)
    return {"workload": "the headline workload in float32", "ms_per_step": dt * 1e3, "elements_per_s": E / dt,
            "nnz_per_s": nnz / dt, "kernel_us": k_s * 1e6, "pipeline_bytes_fp32": gb + sb,
            "frac": (gb + sb) / k_s / PEAK_HBM,
            "error_vs_fp64": {"normwise": (d.norm() / K64.values.norm()).item(),
                              "max_abs_over_max": (d.abs().max() / K64.values.abs().max()).item()}}


def load_traffic():
    """Calibrated HBM bytes per launch from the committed counter summary, or
    None when it was measured on other kernel sources than these."""
    path = os.path.join(ROOT, "profiles", "p1_tet_diffusion_1572864_f64_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    if data.get("source_digest") != source_digest():
        return None
    return data.get("bytes_per_launch", {})


def profile_kernels(reps=5):
    """Launch each assembly kernel of the workload ``reps`` times, plus a plain
    256 MiB device-to-device copy (known bytes: 256 MiB read + 256 MiB
    written) that calibrates the HBM counters, for rocprofv3 runs."""
    mesh, kappa, A, _, _ = build_workload(torch.float64)
    ed = {"kappa": kappa}
    for path in ("rows", "two_stage", "two_stage_rows", "incidence"):
        A.p1_path = path
        for _ in range(reps):
            K = A(element_data=ed)
    # a few PCG iterations on the condensed system (two-launch kernels)
    import tensormesh_amd as tm
    from tensormesh_amd.solve import BatchedPCG
    pts = mesh.points
    bnd = ((pts <= 1e-12) | (pts >= 1.0 - 1e-12)).any(dim=1)
    Ki, fi = tm.Condenser(bnd)(K, torch.ones(mesh.n_points, dtype=torch.float64, device="cuda"))
    BatchedPCG(Ki.row_ptr, Ki.col_indices, Ki.values, [0, Ki.shape[0]], rtol=1e-30).solve(fi, max_iter=64,
                                                                                          graph=False)
    src = torch.empty(256 * 1024 * 1024 // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"profiled": ["p1_rows_dest_kernel", "p1_element_kernel", "reduce_stream_kernel",
                                   "p1_rows_reduce_kernel", "p1_element_inc_kernel", "cg2_vec_kernel",
                                   "cg2_spmv_kernel", "copy 268435456 B"], "reps": reps}))


def profile_hyper(reps=3):
    """Launch the Neo-Hookean residual + tangent assembly a few times per
    precision and reduction path, for rocprofv3 runs."""
    import tensormesh_amd as tm
    from tensormesh_amd.mesh import unit_cube_tet
    m = unit_cube_tet(32, 0.0, 0, order=2).to("cuda")
    model = tm.NeoHookeanModel.from_mesh(m, quadrature_order=2, E=1.0, nu=0.3)
    for dt, mfma, path in ((torch.float64, False, "slots"), (torch.float64, False, "rows"),
                           (torch.float32, False, "slots"), (torch.float32, True, "slots")):
        model.use_mfma = mfma
        model.nh_path = path
        pts = m.points.to(dt)
        u = torch.zeros_like(pts)
        u[:, 0] = 0.1 * pts[:, 0]
        for _ in range(reps):
            model.residual_and_tangent(u, points=pts)
    torch.cuda.synchronize()
    print(json.dumps({"profiled": "neo-hookean tetra10 32^3x6", "reps": reps}))


def profile_forms(reps=3):
    """Launch the energy-form and block-form assemblies a few times each, for
    rocprofv3 counter runs: Neo-Hookean tetra10 32^3x6 (float64 symmetric
    blocks, float32, float32 MFMA tangent, energy + backward), J2 residual +
    tangent on 48^3x6 P1 tets, hexahedral elasticity on 64^3 cells, plus the
    256 MiB calibration copy."""
    import tensormesh_amd as tm
    from tensormesh_amd.mesh import unit_cube_hex, unit_cube_tet
    m = unit_cube_tet(32, 0.0, 0, order=2).to("cuda")
    model = tm.NeoHookeanModel.from_mesh(m, quadrature_order=2, E=1.0, nu=0.3)
    for dt, mfma in ((torch.float64, False), (torch.float32, False), (torch.float32, True)):
        model.use_mfma = mfma
        pts = m.points.to(dt)
        u = torch.zeros_like(pts)
        u[:, 0] = 0.1 * pts[:, 0]
        for _ in range(reps):
            model.residual_and_tangent(u, points=pts)
    model.use_mfma = False
    u = torch.zeros_like(m.points)
    u[:, 0] = 0.1 * m.points[:, 0]
    for _ in range(reps):
        x = u.clone().requires_grad_(True)
        model.energy(x).backward()
    mj = unit_cube_tet(48, 0.1, 0).to("cuda")
    j2 = tm.J2Plasticity.from_mesh(mj, quadrature_order=2, E=200.0, nu=0.3, sig0=0.25, H=1.0)
    P = mj.points
    uj = torch.stack([3e-3 * P[:, 0], -1e-3 * P[:, 1], 5e-4 * P[:, 2]], 1)
    for _ in range(reps):
        j2.residual_and_tangent(uj)
    mh = unit_cube_hex(64, 0.15, 0).to("cuda")
    A = tm.LinearElasticityElementAssembler.from_mesh(mh, quadrature_order=2, E=1.0, nu=0.3)
    for _ in range(reps):
        A()
    src = torch.empty(256 * 1024 * 1024 // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"profiled": ["hyper_wave_kernel (f64, f32, f32 MFMA, energy)", "reduce_stream_kernel (9)",
                                   "j2_wave_kernel", "wave_element_kernel (hexahedron)", "copy 268435456 B"],
                      "reps": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--dtype", choices=["f64", "f32"], default="f64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-grad-step", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the Neo-Hookean and batched legs")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="only launch each assembly kernel a few times (for rocprofv3 counter runs)")
    ap.add_argument("--profile-hyper", action="store_true",
                    help="only launch the Neo-Hookean assembly a few times (for rocprofv3 runs)")
    ap.add_argument("--profile-forms", action="store_true",
                    help="only launch the Neo-Hookean, J2 and hexahedral assemblies a few times (rocprofv3)")
    ap.add_argument("--legs", default="",
                    help="comma-separated side legs to run alone, one JSON line each "
                         "(plasticity, contact, hexahedra, grad_step, facet, neo_hookean, batched)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note, not a measured path)"}))
        sys.exit(2)
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no ROCm GPU visible"}))
        sys.exit(2)

    if args.profile_kernels:
        profile_kernels()
        return
    if args.profile_hyper:
        profile_hyper()
        return
    if args.profile_forms:
        profile_forms()
        return
    if args.legs:
        legs = {"plasticity": plasticity_bench, "contact": contact_bench, "hexahedra": hex_bench,
                "grad_step": grad_step_bench, "fp32": fp32_bench,
                "facet": facet_bench, "neo_hookean": hyperelastic_bench, "batched": batched_bench}
        for name in args.legs.split(","):
            print(json.dumps({name: legs[name]()}), flush=True)
        return
    dtype = torch.float64 if args.dtype == "f64" else torch.float32
    mesh, kappa, A, setup_s, host = build_workload(dtype)
    E = A.elements["tetra"].shape[0]
    N = mesh.n_points
    nnz = A.pattern.nnz
    ed = {"kappa": kappa}

    def timed(path):
        A.p1_path = path
        for _ in range(args.warmup):
            K = A(element_data=ed)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            K = A(element_data=ed)
        torch.cuda.synchronize()
        assert K.nnz == nnz
        return (time.perf_counter() - t0) / args.steps

    dt_two = timed("two_stage")
    dt_two_rows = timed("two_stage_rows")
    dt_inc = timed("incidence")
    dt_fused = timed("rows")
    dt = min(dt_two, dt_two_rows, dt_fused, dt_inc)

    g_s, s_s, s2_s = stage_times(A, mesh, kappa)
    f_s, f_blocks = fused_time(A, mesh, kappa)
    inc_times = incidence_stage_times(A, mesh, kappa)
    s = 8 if dtype == torch.float64 else 4
    gb, sb = algorithmic_bytes(E, N, nnz, s=s, c=s)
    s2b = rows_reduce_bytes(E, N, nnz, s=s)
    fb = fused_bytes(E, N, nnz, s=s, c=s)
    traffic = load_traffic() or {}
    # kernels of the path the API call runs (the fastest one, set as default below)
    kernels = {"fused_rows": [("p1_rows_dest_kernel", fb, f_s)],
               "two_stage_segments": [("p1_element_kernel", gb, g_s), ("reduce_stream_kernel", sb, s_s)],
               "two_stage_rows": [("p1_element_kernel", gb, g_s), ("p1_rows_reduce_kernel", s2b, s2_s)]}
    best = min((("fused_rows", dt_fused), ("two_stage_segments", dt_two), ("two_stage_rows", dt_two_rows)),
               key=lambda x: x[1])[0]
    if dt_inc < min(dt_fused, dt_two, dt_two_rows):
        best = "incidence"
        kernels["incidence"] = [("p1_element_inc_kernel", gb + 4 * 4 * E, inc_times[0]),
                                ("p1_rows_reduce_kernel<inc_order>", s2b - 4 * 4 * E, inc_times[1])]
    name, kb, ks = max(kernels[best], key=lambda k: k[2])
    # Algorithmic bytes per launch: the per-element unique bytes of the two-stage pipeline
    # (stage G + stage S, every array once, int32 indices) times the elements one launch
    # assembles.  The fused kernel does both stages' work in one launch, so its launch is
    # charged the whole pipeline's bytes; a stage kernel is charged its own stage's bytes.
    # The bytes the kernel itself must move (its inputs, index structures and output) are
    # reported beside it as kernel_own_bytes.
    per_launch = gb + sb if best == "fused_rows" else kb
    ach = per_launch / ks
    roof = {"bound": "hbm", "kernel": name, "path": best,
            "achieved": ach / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": ach / PEAK_HBM,
            "traffic": traffic.get(name), "algorithmic_bytes": per_launch,
            "denominator": ("two-stage pipeline unique bytes (stage G + stage S) per launch"
                            if best == "fused_rows" else "the stage's unique bytes per launch"),
            "kernel_own_bytes": kb, "kernel_own_frac": kb / ks / PEAK_HBM, "kernel_us": ks * 1e6}
    pipeline_bw = (gb + sb) / (g_s + s_s)

    out = {
        "metric": "elements assembled/sec + %HBM roofline; fwd+adjoint grad step time vs plain",
        "value": E / dt,
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic (structured jittered Kuhn tetrahedra, seeded log-normal kappa)",
        "config": {"workload": "P1 tet variable-coefficient diffusion stiffness, 64^3x6 = 1,572,864 tets, "
                               "274,625 nodes, 4,018,753 CSR entries, qorder 1",
                   "elements": E, "nodes": N, "nnz": nnz, "parallelism": "single GPU"},
        "nnz_per_s": nnz / dt,
        "paths_ms_per_step": {"fused_rows": dt_fused * 1e3, "two_stage_segments": dt_two * 1e3,
                              "two_stage_rows": dt_two_rows * 1e3, "incidence": dt_inc * 1e3},
        "two_stage": {"ms_per_step": dt_two * 1e3, "stage_G_us": g_s * 1e6, "stage_S_us": s_s * 1e6,
                      "stage_S_rows_us": s2_s * 1e6, "bytes_S_rows": s2b, "stage_S_rows_frac": s2b / s2_s / PEAK_HBM,
                      "fused_kernel_us": f_s * 1e6, "bytes_fused": fb, "fused_frac": fb / f_s / PEAK_HBM,
                      "incidence_G_us": inc_times[0] * 1e6, "incidence_S_us": inc_times[1] * 1e6,
                      "incidence_G_frac": (gb + 16 * E) / inc_times[0] / PEAK_HBM,
                      "incidence_S_frac": (s2b - 16 * E) / inc_times[1] / PEAK_HBM,
                      "bytes_G": gb, "bytes_S": sb,
                      "stage_G_frac": gb / g_s / PEAK_HBM, "stage_S_frac": sb / s_s / PEAK_HBM,
                      "pipeline_GBps": pipeline_bw / 1e9, "pipeline_frac": pipeline_bw / PEAK_HBM,
                      "traffic": None if traffic is None else {k: traffic.get(k) for k in ("G", "S")}},
        "pattern_setup_s": setup_s,
        "roofline": roof,
        # the compulsory floor (conn, coords, kappa in; values out) over the whole call's time:
        # the bytes no assembly path can avoid moving
        "roofline_compulsory_floor": {"bytes": 4 * 4 * E + s * 3 * N + s * E + s * nnz,
                                      "achieved": (4 * 4 * E + s * 3 * N + s * E + s * nnz) / dt / 1e9,
                                      "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                                      "frac": (4 * 4 * E + s * 3 * N + s * E + s * nnz) / dt / PEAK_HBM},
    }
    out["fp32"] = fp32_bench()
    if not args.no_grad_step:
        out["grad_step"] = grad_step_bench()
    if not args.no_extras:
        out["neo_hookean"] = hyperelastic_bench()
        out["batched"] = batched_bench()
        out["facet"] = facet_bench()
        out["plasticity"] = plasticity_bench()
        out["contact"] = contact_bench()
        out["hexahedra"] = hex_bench()
    if not args.no_cpu_baseline:
        cb = cpu_baseline(*host)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        out["speedup_vs_cpu"] = (E / dt) / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
