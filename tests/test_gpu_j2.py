"""J2 plasticity on the HIP kernels (tm_assemble_j2 / tm_j2_update_state)
against the CPU oracle (oracle/fem.py, J2 section).

Tolerances (float64): energy rtol 1e-12 (psi >= 0, no cancellation);
residual |d| <= 1e-11 |ref| + 1e-12 max|ref|; tangent |d| <= 1e-10 |ref| +
1e-12 max|ref| (the kernels use the closed-form Hessian, the oracle
autograd); history update rtol 1e-12 with a floor of 1e-14 max|ref|.
Material in consistent units: E = 200, nu = 0.3, sig0 = 0.25, H = 1.
"""
import math
import sys

import pytest
import torch

import tensormesh_amd as tm
from oracle import fem
from tensormesh_amd.mesh import unit_cube_tet, unit_square_tri

pytestmark = pytest.mark.gpu
DEV = "cuda"
PRM = dict(E=200.0, nu=0.3, sig0=0.25, H=1.0)
# displacement amplitudes that leave a mix of elastic and yielding points
AMP = {"triangle": 6e-4, "tetra": 6e-4, "triangle6": 6e-4, "tetra10": 2.5e-4}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mesh(etype, n=2, seed=2):
    order = 2 if etype in ("tetra10", "triangle6") else 1
    if etype.startswith("tri"):
        return unit_square_tri(n, 0.1, seed, order=order)
    return unit_cube_tet(n, 0.1, seed, order=order)


def _state(E, Q, seed, scale):
    g = torch.Generator().manual_seed(seed)
    ep = scale * torch.randn(E, Q, 3, 3, generator=g, dtype=torch.float64)
    ep = 0.5 * (ep + ep.transpose(-1, -2))
    ep = ep - ep.diagonal(dim1=-2, dim2=-1).sum(-1)[..., None, None] / 3.0 * torch.eye(3, dtype=torch.float64)
    al = scale * torch.rand(E, Q, generator=g, dtype=torch.float64)
    return ep, al


def _displacement(P, amp, seed):
    g = torch.Generator().manual_seed(seed)
    smooth = torch.stack([torch.sin(P[:, 1]), P[:, 0] * P[:, -1], 0.5 * P[:, 0] ** 2][:P.shape[1]], 1)
    return amp * (smooth + 0.3 * torch.randn(P.shape, generator=g, dtype=torch.float64))


def _yield_fraction(m, etype, u, ep, al):
    mu, _ = fem.j2_moduli(PRM["E"], PRM["nu"])
    G, _, _ = fem.j2_grad_u(m.points, m.elements(), etype, 2, u)
    x = fem._j2_strain3(G) - ep
    dev = x - x.diagonal(dim1=-2, dim2=-1).sum(-1)[..., None, None] / 3 * torch.eye(3, dtype=torch.float64)
    f = 2 * mu * torch.linalg.norm(dev, dim=(-2, -1)) - math.sqrt(2 / 3) * (PRM["sig0"] + PRM["H"] * al)
    return (f > 0).double().mean().item()


def _within(a, b, rtol, floor):
    a, b = a.detach().double().cpu(), b.detach().double().cpu()
    err = (a - b).abs()
    ok = bool((err <= rtol * b.abs() + floor * b.abs().max()).all())
    return ok, err.max().item() / max(b.abs().max().item(), 1e-300)


def _model(m, dtype=torch.float64):
    md = tm.Mesh(m.points.to(dtype), {m.default_element_type: m.elements()}).to(DEV)
    return tm.J2Plasticity.from_mesh(md, quadrature_order=2, **PRM)


@pytest.mark.parametrize("etype", ["triangle", "tetra", "triangle6", "tetra10"])
def test_j2_energy_residual_tangent_against_oracle(etype):
    say(f"case j2 energy/residual/tangent {etype}")
    m = _mesh(etype)
    C = m.elements()
    Q = fem.quadrature(etype, 2)[0].shape[0]
    ep, al = _state(C.shape[0], Q, 11, 1e-4)
    u = _displacement(m.points, AMP[etype] * (1.2 if etype == "tetra10" else 1.0), 12)
    frac = _yield_fraction(m, etype, u, ep, al)
    say(f"  yielded fraction {frac:.2f}")
    assert 0.1 < frac < 0.9
    D = m.points.shape[1]
    e_ref = fem.j2_energy(m.points, C, etype, 2, u, ep, al, **PRM)
    res_e, tan_e = fem.j2_element_derivatives(m.points, C, etype, 2, u, ep, al, **PRM)
    res = torch.zeros(m.n_points, D, dtype=torch.float64).index_add(0, C.reshape(-1), res_e.reshape(-1, D))
    row, col, eids = fem.pattern(C.numpy(), m.n_points)
    tan = fem._index_add(tan_e, torch.from_numpy(eids), row.size)
    v, r, c = fem.block_coo(tan, torch.from_numpy(row), torch.from_numpy(col))

    model = _model(m)
    ud = u.to(DEV).requires_grad_(True)
    hist = {"eps_p_n": {etype: ep.to(DEV)}, "alpha_n": {etype: al.to(DEV)}}
    energy = model.energy(point_data={"displacement": ud}, element_data=hist)
    assert energy.device.type == "cuda" and energy.dim() == 0
    assert abs(energy.item() - e_ref.item()) <= 1e-12 * abs(e_ref.item())
    energy.backward()
    ok, err = _within(ud.grad.reshape(-1), res.reshape(-1), 1e-11, 1e-12)
    assert ok, f"energy backward vs oracle residual: {err:.2e}"
    R, K = model.residual_and_tangent(u.to(DEV), ep.to(DEV), al.to(DEV))
    ok, err = _within(R, res.reshape(-1), 1e-11, 1e-12)
    assert ok, f"residual: {err:.2e}"
    assert torch.equal(K.row.cpu(), r) and torch.equal(K.col.cpu(), c)
    ok, err = _within(K.values, v, 1e-10, 1e-12)
    assert ok, f"tangent: {err:.2e}"
    say(f"  energy {energy.item():.6e}  tangent max rel err {err:.2e}")


@pytest.mark.parametrize("etype", ["triangle", "tetra10"])
def test_j2_update_state_against_oracle(etype):
    say(f"case j2 update_state {etype}")
    m = _mesh(etype, n=3, seed=4)
    C = m.elements()
    Q = fem.quadrature(etype, 2)[0].shape[0]
    ep, al = _state(C.shape[0], Q, 21, 1e-4)
    u = _displacement(m.points, AMP[etype], 22)
    frac = _yield_fraction(m, etype, u, ep, al)
    assert 0.1 < frac < 0.9, frac
    ep1, al1 = fem.j2_update(m.points, C, etype, 2, u, ep, al, **PRM)
    model = _model(m)
    model.history[etype]["eps_p"].copy_(ep)
    model.history[etype]["alpha"].copy_(al)
    model.update_state(u.to(DEV))
    h = model.history[etype]
    assert h["eps_p"].device.type == "cuda"
    ok, err = _within(h["eps_p"], ep1, 1e-12, 1e-14)
    assert ok, f"eps_p: {err:.2e}"
    ok, err = _within(h["alpha"], al1, 1e-12, 1e-14)
    assert ok, f"alpha: {err:.2e}"
    # the default history feeds residual() after the update
    R = model.residual(u.to(DEV))
    res_e, _ = fem.j2_element_derivatives(m.points, C, etype, 2, u, ep1, al1, **PRM)
    D = m.points.shape[1]
    res = torch.zeros(m.n_points, D, dtype=torch.float64).index_add(0, C.reshape(-1), res_e.reshape(-1, D))
    ok, err = _within(R, res.reshape(-1), 1e-10, 1e-11)
    assert ok, f"residual after update: {err:.2e}"


@pytest.mark.parametrize("etype", ["triangle6", "tetra"])
def test_j2_zero_state_tangent_is_linear_elasticity(etype):
    """At u = 0 with no plastic strain the deviator vanishes; the kernels use
    the elastic limit there, which is the linear-elasticity stiffness (the
    native LinearElasticityElementAssembler).  torch autograd's Hessian is
    degenerate at this point, so the comparison is with the linear form."""
    say(f"case j2 zero-state tangent {etype}")
    m = _mesh(etype, n=3, seed=5)
    model = _model(m)
    u0 = torch.zeros_like(m.points).to(DEV)
    R, K = model.residual_and_tangent(u0)
    assert R.abs().max().item() == 0.0
    lin = tm.LinearElasticityElementAssembler.from_mesh(m.to(DEV), quadrature_order=2,
                                                       E=PRM["E"], nu=PRM["nu"])()
    assert torch.equal(K.row, lin.row) and torch.equal(K.col, lin.col)
    ok, err = _within(K.values, lin.values, 1e-12, 1e-14)
    assert ok, err


def test_j2_float32_against_float64():
    say("case j2 tetra10 fp32 vs fp64")
    m = _mesh("tetra10", n=3, seed=6)
    C = m.elements()
    Q = fem.quadrature("tetra10", 2)[0].shape[0]
    ep, al = _state(C.shape[0], Q, 31, 1e-4)
    u = _displacement(m.points, AMP["tetra10"], 32)
    m64, m32 = _model(m), _model(m, torch.float32)
    R64, K64 = m64.residual_and_tangent(u.to(DEV), ep.to(DEV), al.to(DEV))
    R32, K32 = m32.residual_and_tangent(u.float().to(DEV), ep.float().to(DEV), al.float().to(DEV))
    assert K32.values.dtype == torch.float32
    dk = (K32.values.double() - K64.values).norm() / K64.values.norm()
    dr = (R32.double() - R64).norm() / R64.norm()
    hist = {"eps_p_n": ep.to(DEV), "alpha_n": al.to(DEV)}
    e64 = m64.energy(point_data={"displacement": u.to(DEV)}, element_data=hist)
    hist32 = {"eps_p_n": ep.float().to(DEV), "alpha_n": al.float().to(DEV)}
    e32 = m32.energy(point_data={"displacement": u.float().to(DEV)}, element_data=hist32)
    de = abs(e32.item() - e64.item()) / abs(e64.item())
    say(f"  fp32 vs fp64: tangent {dk:.2e} residual {dr:.2e} energy {de:.2e}")
    # float32 inputs (u, eps_p) carry 6e-8 relative error; the residual is a
    # difference of stresses of size E * strain
    assert dk < 1e-5 and dr < 1e-4 and de < 1e-5


def test_j2_newton_load_step_converges_quadratically():
    """A uniaxial load step on the unit cube (x = 0 held, x = 1 pulled to
    0.3 % strain, past yield; the starting guess moves only the pulled
    face), solved by Newton with the consistent tangent and a backtracking
    line search on the native energy: the free residual falls below 1e-10 of
    its start within 12 iterations, the last ratios quadratic, and
    update_state then records plastic flow."""
    say("case j2 newton load step")
    m = unit_cube_tet(3, 0.0, 1)
    model = _model(m)
    P = m.points
    fixed = torch.zeros_like(P, dtype=torch.bool)
    fixed[P[:, 0] < 1e-9, 0] = True
    fixed[P[:, 0] > 1 - 1e-9, 0] = True
    fixed[(P.norm(dim=1) < 1e-9), :] = True
    fixed[(P[:, 0] < 1e-9) & (P[:, 1] > 1 - 1e-9) & (P[:, 2] < 1e-9), 2] = True
    u = torch.zeros_like(P)
    u[P[:, 0] > 1 - 1e-9, 0] = 3e-3
    u = u.to(DEV)
    free = (~fixed).reshape(-1).to(DEV)
    h = model.history["tetra"]
    hist = {"eps_p_n": h["eps_p"], "alpha_n": h["alpha"]}
    energy = lambda x: model.energy(point_data={"displacement": x}, element_data=hist).item()
    norms = []
    for it in range(14):
        R, K = model.residual_and_tangent(u)
        norms.append(R[free].norm().item())
        if norms[-1] < 1e-10 * norms[0]:
            break
        Kd = K.to_dense()[free][:, free]
        du = torch.zeros_like(R)
        du[free] = torch.linalg.solve(Kd.cpu(), -R[free].cpu()).to(DEV)
        du = du.reshape(P.shape)
        e0, slope, t = energy(u), (R.reshape(P.shape) * du).sum().item(), 1.0
        while energy(u + t * du) > e0 + 1e-4 * t * slope and t > 1e-4:
            t *= 0.5
        u = u + t * du
    say("  residual norms " + " ".join(f"{x:.2e}" for x in norms))
    assert norms[-1] < 1e-10 * norms[0] and len(norms) <= 13
    assert norms[-2] < 1e-2 * norms[-3]
    model.update_state(u)
    al = model.history["tetra"]["alpha"]
    assert (al > 0).any() and torch.isfinite(al).all()
