"""Parity at the full sizes of the timed workloads.

* Neo-Hookean tetra10 on the 32^3 x 6 Kuhn cube (196,608 elements, 823,875
  dofs) at u = 0.1 x: residual and tangent block rows of 1,000 seeded nodes
  against the oracle's element derivatives of exactly the elements touching
  them (float64: |d| <= 1e-11 |ref| + 1e-13 max|ref|), the tangent's symmetry,
  the tangent at u = 0 against linear elasticity on the same mesh, and the
  float32 VALU and MFMA tangents against the float64 oracle (normwise 1e-6).
* Adjoint gradient on the 512^2 x 2 triangle mesh (524,288 elements): the
  hand-written assembly adjoint against the oracle's autograd at full size,
  and the whole assemble -> condense -> solve -> loss gradient against central
  finite differences on 64 seeded directions (h = 1e-6 to the loss's float64
  rounding floor, the fourth-order stencil at h = 1e-4 to 1e-6).
* Batched solves: all 256 uneven meshes in one assembly, eight of them
  against the per-mesh oracle, every system solved to 1e-8.
"""
import sys

import numpy as np
import pytest
import torch

import tensormesh_amd as tm
from oracle import fem
from tensormesh_amd.mesh import lognormal_coefficient, unit_cube_tet, unit_square_tri

pytestmark = pytest.mark.gpu
DEV = "cuda"


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# ------------------------------------------------------------- Neo-Hookean
_NH = {}


def _nh_setup():
    if not _NH:
        m = unit_cube_tet(32, 0.0, 0, order=2)
        u = torch.zeros_like(m.points)
        u[:, 0] = 0.1 * m.points[:, 0]
        _NH.update(mesh=m, u=u)
    return _NH["mesh"], _NH["u"]


def _sampled_rows_oracle(m, u, n_rows=1000, seed=5, E=1.0, nu=0.3):
    """Oracle residual rows and tangent block rows of n_rows seeded nodes:
    element derivatives of every element touching them, summed in ascending
    element order (the reference's index_add_ order)."""
    conn = m.elements().numpy()
    N = m.n_points
    rows = np.sort(np.random.default_rng(seed).choice(N, n_rows, replace=False))
    touch = np.flatnonzero(np.isin(conn, rows).any(1))            # ascending element ids
    sub = torch.from_numpy(conn[touch])
    res_e, tan_e = fem.neohookean_element_tangent(m.points, sub, "tetra10", 2, u, E=E, nu=nu)
    B = conn.shape[1]
    blocks, res = {}, {}
    is_row = np.zeros(N, bool)
    is_row[rows] = True
    for k, e in enumerate(touch):
        for i in range(B):
            r = conn[e, i]
            if not is_row[r]:
                continue
            res[r] = res.get(r, 0.0) + res_e[k, i]
            for j in range(B):
                key = (r, conn[e, j])
                blocks[key] = blocks[key] + tan_e[k, i, j] if key in blocks else tan_e[k, i, j].clone()
    return rows, res, blocks


def _compare_rows(pat, vals, rows, blocks, D=3):
    rp = pat.row_ptr.cpu().numpy()
    ci = pat.col_idx.cpu().numpy()
    vals = vals.reshape(-1, D, D).cpu().double()
    got, want = [], []
    for r in rows:
        cols = ci[rp[r]:rp[r + 1]]
        for t, c in zip(range(rp[r], rp[r + 1]), cols):
            got.append(vals[t])
            want.append(blocks[(r, c)])
    return torch.stack(got), torch.stack(want)


def test_cfg5_neohookean_full_size_sampled_rows_fp64():
    say("case cfg5 Neo-Hookean 32^3x6 tetra10 fp64, 1000 sampled block rows vs oracle")
    m, u = _nh_setup()
    model = tm.NeoHookeanModel.from_mesh(m.to(DEV), quadrature_order=2, E=1.0, nu=0.3)
    R, K = model.residual_and_tangent(u.to(DEV))
    pat = model.pattern
    assert pat.nnz * 9 == 69108489 and K.nnz == 69108489
    rows, res, blocks = _sampled_rows_oracle(m, u)
    got, want = _compare_rows(pat, K.values, rows, blocks)
    err = (got - want).abs()
    scale = want.abs().max()
    say(f"  tangent rows: {got.shape[0]} blocks, max|d|/max|ref| {(err.max() / scale).item():.2e}")
    assert bool((err <= 1e-11 * want.abs() + 1e-13 * scale).all())
    Rr = R.reshape(-1, 3).cpu()[torch.from_numpy(rows)]
    Rw = torch.stack([res[r] for r in rows])
    rerr = (Rr - Rw).abs()
    say(f"  residual rows: max|d|/max|ref| {(rerr.max() / Rw.abs().max()).item():.2e}")
    assert bool((rerr <= 1e-11 * Rw.abs() + 1e-13 * Rw.abs().max()).all())


def test_cfg5_tangent_symmetric_and_linear_at_rest():
    say("case cfg5 tangent symmetry (u = 0.1x) and u = 0 vs linear elasticity, full size")
    m, u = _nh_setup()
    md = m.to(DEV)
    model = tm.NeoHookeanModel.from_mesh(md, quadrature_order=2, E=1.0, nu=0.3)
    for sym in (None, False):
        model.symmetric_slots = sym
        K = model.tangent(u.to(DEV))
        n = K.shape[0]
        key = K.row * n + K.col
        tkey = K.col * n + K.row
        order = torch.argsort(key)
        pos = order[torch.searchsorted(key[order], tkey)]
        assert torch.equal(key[pos], tkey)
        asym = (K.values - K.values[pos]).abs().max().item() / K.values.abs().max().item()
        say(f"  symmetric_slots={sym}: max|K - K^T| / max|K| = {asym:.2e}")
        assert asym <= 1e-14
        del key, tkey, order, pos
    model.symmetric_slots = None
    K0 = model.tangent(torch.zeros_like(md.points))
    L = tm.LinearElasticityElementAssembler.from_mesh(md, quadrature_order=2, E=1.0, nu=0.3)()
    assert torch.equal(K0.row, L.row) and torch.equal(K0.col, L.col)
    d = (K0.values - L.values).abs()
    scale = L.values.abs().max()
    say(f"  u = 0 vs linear elasticity: max|d|/max|K| {(d.max() / scale).item():.2e}")
    assert bool((d <= 1e-12 * L.values.abs() + 1e-14 * scale).all())


@pytest.mark.parametrize("mfma", [False, True])
def test_cfg5_fp32_tangents_against_fp64_oracle(mfma):
    say(f"case cfg5 fp32 tangent (mfma={mfma}) vs fp64 oracle, 1000 sampled block rows")
    m, _ = _nh_setup()
    p32 = m.points.float()
    u32 = torch.zeros_like(p32)
    u32[:, 0] = 0.1 * p32[:, 0]
    m64 = tm.Mesh(p32.double(), {"tetra10": m.elements()})
    rows, res, blocks = _sampled_rows_oracle(m64, u32.double(), seed=6)
    md = tm.Mesh(p32, {"tetra10": m.elements()}).to(DEV)
    model = tm.NeoHookeanModel.from_mesh(md, quadrature_order=2, E=1.0, nu=0.3)
    model.use_mfma = mfma
    R, K = model.residual_and_tangent(u32.to(DEV))
    assert K.dtype == torch.float32
    got, want = _compare_rows(model.pattern, K.values, rows, blocks)
    normwise = ((got - want).norm() / want.norm()).item()
    maxrel = ((got - want).abs().max() / want.abs().max()).item()
    say(f"  fp32 tangent rows: normwise {normwise:.2e}, max|d|/max|ref| {maxrel:.2e}")
    assert normwise <= 1e-6 and maxrel <= 1e-6
    Rr = R.reshape(-1, 3).double().cpu()[torch.from_numpy(rows)]
    Rw = torch.stack([res[r] for r in rows])
    rn = ((Rr - Rw).norm() / Rw.norm()).item()
    say(f"  fp32 residual rows: normwise {rn:.2e}")
    assert rn <= 1e-5


# --------------------------------------------------------- adjoint, cfg3
_G = {}


def _cfg3():
    if not _G:
        m = unit_square_tri(512, 0.2, 0)
        E = m.elements().shape[0]
        _G.update(mesh=m, kap=torch.from_numpy(lognormal_coefficient(E, 1)),
                  kap_t=torch.from_numpy(lognormal_coefficient(E, 2)))
    return _G["mesh"], _G["kap"], _G["kap_t"]


def test_cfg3_assembly_adjoint_full_size_against_oracle_autograd():
    say("case cfg3 512^2x2 assembly adjoint vs oracle autograd (full size)")
    m, kap, _ = _cfg3()
    P = m.points.clone().requires_grad_(True)
    Kp = kap.clone().requires_grad_(True)
    v, r, c, _ = fem.assemble_bilinear(P, m.elements(), "triangle", 2, "diffusion_elem", kappa=Kp)
    w = torch.randn(v.shape, generator=torch.Generator().manual_seed(3), dtype=torch.float64)
    (v * w).sum().backward()
    md = m.to(DEV)
    Pd = md.points.clone().requires_grad_(True)
    Kd = kap.to(DEV).clone().requires_grad_(True)
    M = tm.DiffusionElementAssembler.from_mesh(md)(Pd, element_data={"kappa": Kd})
    assert "P1Assemble" in type(M.values.grad_fn).__name__
    (M.values * w.to(DEV)).sum().backward()
    ex = ((Pd.grad.cpu() - P.grad).abs().max() / P.grad.abs().max()).item()
    ek = ((Kd.grad.cpu() - Kp.grad).abs().max() / Kp.grad.abs().max()).item()
    say(f"  dL/dx rel err {ex:.2e}, dL/dkappa rel err {ek:.2e}")
    assert ex < 1e-11 and ek < 1e-11


def test_cfg3_gradient_step_finite_differences_full_size():
    """The whole assemble -> condense -> solve -> loss gradient on 512^2 x 2
    against finite differences on 64 seeded directions (kappa and interior
    node coordinates together).

    The loss itself carries a rounding floor: cond(K) amplifies float64
    rounding to about 2e-11 relative in L (measured with an exact sparse LU
    of the same systems on the host: central differences at h = 1e-6 then
    disagree with the exact gradient by up to 6e-4 per direction and 2e-5
    normwise).  So the central differences at h = 1e-6 are held to that
    floor (5e-5 normwise), and the fourth-order central stencil at h = 1e-4
    (truncation O(h^4), rounding noise / 1e-4) to 1e-6 normwise.  Every
    solve takes one refinement step, so the solver's stopping error stays
    below the rounding floor."""
    say("case cfg3 512^2x2 gradient vs central finite differences, 64 directions")
    m0, kap0, kapt0 = _cfg3()
    m = m0.to(DEV)
    kap, kap_t = kap0.to(DEV), kapt0.to(DEV)
    E = kap.numel()
    A = tm.DiffusionElementAssembler.from_mesh(m)
    cond = tm.Condenser(m.boundary_mask)
    x, y = m.points[:, 0], m.points[:, 1]
    f = 2 * np.pi ** 2 * torch.sin(np.pi * x) * torch.sin(np.pi * y)
    b = tm.SourceNodeAssembler.from_mesh(m)(point_data={"f": f})
    kw = dict(rtol=1e-13, max_iter=20000, refine=1)

    def loss(points, kappa):
        K = A(points, element_data={"kappa": kappa})
        Ki, fi = cond(K, b)
        u = Ki.solve(fi, **kw)
        return ((u - u_t) ** 2).sum()

    with torch.no_grad():
        Kt = A(element_data={"kappa": kap_t})
        Kti, fti = cond(Kt, b)
        u_t = Kti.solve(fti, **kw)
        say(f"  target solve: {Kti.last_solve_info[0]} iterations, true relres {Kti.last_solve_info[1]:.2e}")
    kr = kap.clone().requires_grad_(True)
    pr = m.points.clone().requires_grad_(True)
    loss(pr, kr).backward()
    interior = (~m.boundary_mask).to(torch.float64)[:, None]
    g = torch.Generator().manual_seed(11)
    c2, c4, ans = [], [], []
    for t in range(64):
        dk = torch.randn(E, generator=g, dtype=torch.float64).to(DEV)
        dx = torch.randn(m.points.shape, generator=g, dtype=torch.float64).to(DEV) * interior * 1e-3
        with torch.no_grad():
            at = lambda s: loss(m.points + s * dx, kap + s * dk).item()
            h = 1e-6
            c2.append((at(h) - at(-h)) / (2 * h))
            h = 1e-4
            c4.append((at(-2 * h) - 8 * at(-h) + 8 * at(h) - at(2 * h)) / (12 * h))
        ans.append(((kr.grad * dk).sum() + (pr.grad * dx).sum()).item())
        if t % 16 == 15:
            say(f"  {t + 1} directions done")
    c2, c4, ans = np.array(c2), np.array(c4), np.array(ans)
    w2 = np.abs(c2 - ans).max() / np.abs(ans).max()
    w4 = np.abs(c4 - ans).max() / np.abs(ans).max()
    say(f"  central h=1e-6: max |fd - adjoint| / max |adjoint| = {w2:.2e} "
        f"(per-direction median {np.median(np.abs(c2 - ans) / np.abs(ans)):.2e})")
    say(f"  fourth-order h=1e-4: max |fd - adjoint| / max |adjoint| = {w4:.2e} "
        f"(per-direction median {np.median(np.abs(c4 - ans) / np.abs(ans)):.2e})")
    assert w2 <= 5e-5
    assert w4 <= 1e-6


# ------------------------------------------------------------ batched, cfg4
def test_cfg4_all_256_meshes_sampled_against_oracle():
    from tensormesh_amd.batched import BatchedDiffusion
    say("case cfg4 256 uneven meshes: one assembly, 8 sampled vs per-mesh oracle, every system to 1e-8")
    rng = np.random.default_rng(0)
    sizes = rng.integers(16, 129, size=256)
    meshes = [unit_square_tri(int(nk), 0.2, 1000 + k) for k, nk in enumerate(sizes)]
    kap = torch.cat([torch.from_numpy(lognormal_coefficient(mk.elements().shape[0], 2000 + k))
                     for k, mk in enumerate(meshes)])
    f = torch.cat([torch.sin(np.pi * mk.points[:, 0]) * torch.sin(np.pi * mk.points[:, 1]) * (1 + k % 3)
                   for k, mk in enumerate(meshes)])
    bp = BatchedDiffusion([mk.to(DEV) for mk in meshes])
    Ki, fi = bp.assemble(kap.to(DEV), f.to(DEV), torch.float64)
    sample = np.sort(np.random.default_rng(1).choice(256, 8, replace=False))
    inner_off = bp.sys_rows
    kvals = Ki.values.cpu()
    krow = Ki.row.cpu()
    fi_c = fi.cpu()
    for k in sample:
        mk = meshes[k]
        kk = kap[bp.elem_off[k]:bp.elem_off[k + 1]]
        ff = f[bp.node_off[k]:bp.node_off[k + 1]]
        v, r, c, _ = fem.assemble_bilinear(mk.points, mk.elements(), "triangle", 2, "diffusion_elem", kappa=kk)
        bk = fem.assemble_linear(mk.points, mk.elements(), "triangle", 2, ff)
        cd = fem.condense(v, r, c, mk.n_points, mk.boundary_mask, rhs=bk)
        sel = (krow >= inner_off[k]) & (krow < inner_off[k + 1])
        got = kvals[sel]
        want = cd["K_ii"][0]
        assert got.numel() == want.numel()
        e1 = ((got - want).abs().max() / want.abs().max()).item()
        e2 = ((fi_c[inner_off[k]:inner_off[k + 1]] - cd["f_i"]).abs().max() / cd["f_i"].abs().max()).item()
        say(f"  mesh {k} (N={sizes[k]}): K_ii {e1:.1e}, f_i {e2:.1e}")
        assert e1 < 1e-13 and e2 < 1e-13
    u, st = bp.solve(kap.to(DEV), f.to(DEV), rtol=1e-8, mixed=True)
    rel = st["relres"]
    say(f"  mixed-precision solve: outer {st['outer']}, max relres {rel.max().item():.2e}")
    assert rel.numel() == 256 and bool((rel <= 1e-8).all())
