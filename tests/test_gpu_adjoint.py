"""Hand-written adjoint of the P1 assembly vs. the oracle's autograd (the
reference's own route: autograd through einsum / inverse / det / index_add_)
and vs. central finite differences of a full assemble -> condense -> solve ->
loss chain.  Tolerances: 1e-11 relative against autograd (float64); finite
differences with h = 1e-6 agree to 1e-6 relative."""
import sys

import numpy as np
import pytest
import torch

import tensormesh_amd as tm
from oracle import fem
from tensormesh_amd.mesh import lognormal_coefficient, unit_cube_tet, unit_square_tri

pytestmark = pytest.mark.gpu
DEV = "cuda"


def say(m):
    print(m, flush=True)


def _mesh(kind):
    return unit_square_tri(6, 0.2, 5) if kind == "tri" else unit_cube_tet(2, 0.12, 5)


@pytest.mark.parametrize("kind", ["tri", "tet"])
@pytest.mark.parametrize("form", ["laplace", "diffusion_elem", "diffusion_node", "mass"])
def test_adjoint_matches_oracle_autograd(kind, form):
    say(f"case adjoint {kind} {form}")
    m = _mesh(kind)
    etype = m.default_element_type
    E = m.elements().shape[0]
    g = torch.Generator().manual_seed(0)
    kap = None
    if form == "diffusion_elem":
        kap = torch.from_numpy(lognormal_coefficient(E, 2))
    elif form == "diffusion_node":
        kap = torch.from_numpy(lognormal_coefficient(m.n_points, 2))
    # oracle autograd
    P = m.points.clone().requires_grad_(True)
    K = kap.clone().requires_grad_(True) if kap is not None else None
    v, r, c, _ = fem.assemble_bilinear(P, m.elements(), etype, 2, form, kappa=K)
    w = torch.randn(v.shape, generator=g, dtype=torch.float64)
    (v * w).sum().backward()
    # native
    md = m.to(DEV)
    Pd = md.points.clone().requires_grad_(True)
    Kd = kap.to(DEV).clone().requires_grad_(True) if kap is not None else None
    if form == "laplace":
        A = tm.LaplaceElementAssembler.from_mesh(md)
        M = A(Pd)
    elif form == "mass":
        A = tm.MassElementAssembler.from_mesh(md)
        M = A(Pd)
    elif form == "diffusion_elem":
        A = tm.DiffusionElementAssembler.from_mesh(md)
        M = A(Pd, element_data={"kappa": Kd})
    else:
        A = tm.DiffusionElementAssembler.from_mesh(md)
        M = A(Pd, point_data={"kappa": Kd})
    assert M.values.grad_fn is not None and "P1Assemble" in type(M.values.grad_fn).__name__
    (M.values * w.to(DEV)).sum().backward()
    gx, gx_ref = Pd.grad.cpu(), P.grad
    err = ((gx - gx_ref).abs().max() / gx_ref.abs().max()).item()
    say(f"  dL/dx rel err {err:.2e}")
    assert err < 1e-11
    if kap is not None:
        gk, gk_ref = Kd.grad.cpu(), K.grad
        errk = ((gk - gk_ref).abs().max() / gk_ref.abs().max()).item()
        say(f"  dL/dkappa rel err {errk:.2e}")
        assert errk < 1e-11


def _loss_chain(points, kappa, A, cond, b, u_target):
    # one refinement step takes the solve from the CG stopping error
    # (up to cond * rtol) to the rounding level of the residual, so the
    # difference quotient sees the loss, not where the iteration stopped
    K = A(points, element_data={"kappa": kappa})
    Ki, fi = cond(K, b)
    u = Ki.solve(fi, rtol=1e-13, refine=1)
    return ((u - u_target) ** 2).sum()


def test_gradient_step_finite_differences():
    say("case adjoint FD on 24x24 tri diffusion")
    m = unit_square_tri(24, 0.2, 0).to(DEV)
    E = m.elements().shape[0]
    kap = torch.from_numpy(lognormal_coefficient(E, 1)).to(DEV)
    kap_t = torch.from_numpy(lognormal_coefficient(E, 2)).to(DEV)
    A = tm.DiffusionElementAssembler.from_mesh(m)
    cond = tm.Condenser(m.boundary_mask)
    x, y = m.points[:, 0], m.points[:, 1]
    f = 2 * np.pi ** 2 * torch.sin(np.pi * x) * torch.sin(np.pi * y)
    b = tm.SourceNodeAssembler.from_mesh(m)(point_data={"f": f})
    with torch.no_grad():
        Kt = A(element_data={"kappa": kap_t})
        Kti, fti = cond(Kt, b)
        u_target = Kti.solve(fti, rtol=1e-13, refine=1)
    kr = kap.clone().requires_grad_(True)
    pr = m.points.clone().requires_grad_(True)
    L = _loss_chain(pr, kr, A, cond, b, u_target)
    L.backward()
    interior = (~m.boundary_mask).to(torch.float64)[:, None]
    g = torch.Generator(device="cpu").manual_seed(7)
    h = 1e-6
    fds, ans = [], []
    for t in range(8):
        dk = torch.randn(E, generator=g, dtype=torch.float64).to(DEV)
        dx = (torch.randn(m.points.shape, generator=g, dtype=torch.float64).to(DEV) * interior) * 1e-2
        with torch.no_grad():
            lp = _loss_chain(m.points + h * dx, kap + h * dk, A, cond, b, u_target)
            lm = _loss_chain(m.points - h * dx, kap - h * dk, A, cond, b, u_target)
        fds.append(((lp - lm) / (2 * h)).item())
        ans.append(((kr.grad * dk).sum() + (pr.grad * dx).sum()).item())
    fds, ans = np.array(fds), np.array(ans)
    say("  fd " + " ".join(f"{v:+.6e}" for v in fds))
    say("  an " + " ".join(f"{v:+.6e}" for v in ans))
    worst = np.abs(fds - ans).max() / np.abs(fds).max()
    say(f"  FD max |fd - adjoint| / max |fd| over 8 directions: {worst:.2e}")
    assert worst < 1e-6
