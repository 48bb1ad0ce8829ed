"""The order-7 simplex rules (13-point triangle, 24-point tetrahedron,
quadrature.py:96-160) through the HIP kernels, against the oracle at the
same rule.  A tetrahedron rule of 24 points exceeds the 16 that the
four-wave Neo-Hookean and J2 kernels stage per element, so those launches
take the one-wave variant with rules of up to 32 points; the 13-point
triangle rule stays on the four-wave kernels.  Tolerances as the order-2
tests: residuals |d| <= 1e-11 |ref| + 1e-12 max|ref|, tangents and element
matrices 1e-10 / 1e-12, energies 1e-12 relative."""
import sys

import pytest
import torch

import tensormesh_amd as tm
from oracle import fem
from tensormesh_amd.mesh import lognormal_coefficient, unit_cube_tet, unit_square_tri

pytestmark = pytest.mark.gpu
DEV = "cuda"
Q7 = 7


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mesh(etype, seed=2):
    order = 2 if etype in ("tetra10", "triangle6") else 1
    if etype.startswith("tri"):
        return unit_square_tri(3, 0.1, seed, order=order)
    return unit_cube_tet(2, 0.1, seed, order=order)


def _within(a, b, rtol, floor):
    a, b = a.detach().double().cpu(), b.detach().double().cpu()
    err = (a - b).abs()
    ok = bool((err <= rtol * b.abs() + floor * b.abs().max()).all())
    return ok, err.max().item() / max(b.abs().max().item(), 1e-300)


def _smooth(P, amp, seed):
    g = torch.Generator().manual_seed(seed)
    s = torch.stack([torch.sin(P[:, 1]), P[:, 0] * P[:, -1], 0.5 * P[:, 0] ** 2][:P.shape[1]], 1)
    return amp * s + 0.04 * amp * torch.randn(P.shape, generator=g, dtype=torch.float64)


def test_order7_rule_sizes():
    assert fem.quadrature("triangle", Q7)[0].numel() == 13
    assert fem.quadrature("tetra", Q7)[0].numel() == 24


@pytest.mark.parametrize("etype", ["tetra10", "tetra", "triangle6"])
def test_neohookean_order7_against_oracle(etype):
    say(f"case neo-hookean {etype} quadrature order 7")
    m = _mesh(etype)
    u = _smooth(m.points, 0.05, 5)
    C = m.elements()
    D = m.points.shape[1]
    res_e, tan_e = fem.neohookean_element_tangent(m.points, C, etype, Q7, u, E=1.0, nu=0.3)
    res = torch.zeros(m.n_points, D, dtype=torch.float64).index_add(0, C.reshape(-1), res_e.reshape(-1, D))
    row, col, eids = fem.pattern(C.numpy(), m.n_points)
    tan = fem._index_add(tan_e, torch.from_numpy(eids), row.size)
    v, r, c = fem.block_coo(tan, torch.from_numpy(row), torch.from_numpy(col))
    model = tm.NeoHookeanModel.from_mesh(m.to(DEV), quadrature_order=Q7, E=1.0, nu=0.3)
    R, K = model.residual_and_tangent(u.to(DEV))
    ok, err = _within(R, res.reshape(-1), 1e-11, 1e-12)
    assert ok, f"residual {err:.2e}"
    assert torch.equal(K.row.cpu(), r) and torch.equal(K.col.cpu(), c)
    ok, err = _within(K.values, v, 1e-10, 1e-12)
    assert ok, f"tangent {err:.2e}"
    e_ref = fem.neohookean_energy(m.points, C, etype, Q7, u, E=1.0, nu=0.3)
    e = model.energy(u.to(DEV))
    assert abs(e.item() - e_ref.item()) <= 1e-12 * abs(e_ref.item())
    say(f"  tangent max rel err {err:.2e}")


@pytest.mark.parametrize("etype,form", [("tetra10", "laplace"), ("tetra10", "mass"), ("triangle6", "laplace"),
                                        ("tetra", "linear_elasticity"), ("tetra", "diffusion_node")])
def test_bilinear_order7_against_oracle(etype, form):
    say(f"case {etype} {form} quadrature order 7")
    m = _mesh(etype, seed=11)
    kap = torch.from_numpy(lognormal_coefficient(m.n_points, 3))
    ref = fem.assemble_bilinear(m.points, m.elements(), etype, Q7, form, kappa=kap, E=1.0, nu=0.3)
    md = m.to(DEV)
    if form == "laplace":
        K = tm.LaplaceElementAssembler.from_mesh(md, quadrature_order=Q7)()
    elif form == "mass":
        K = tm.MassElementAssembler.from_mesh(md, quadrature_order=Q7)()
    elif form == "diffusion_node":
        K = tm.DiffusionElementAssembler.from_mesh(md, quadrature_order=Q7)(point_data={"kappa": kap.to(DEV)})
    else:
        K = tm.LinearElasticityElementAssembler.from_mesh(md, quadrature_order=Q7, E=1.0, nu=0.3)()
    assert torch.equal(K.row.cpu(), ref[1]) and torch.equal(K.col.cpu(), ref[2])
    ok, err = _within(K.values, ref[0], 1e-12, 1e-14)
    assert ok, f"values {err:.2e}"


@pytest.mark.parametrize("etype", ["tetra", "tetra10"])
def test_j2_order7_against_oracle(etype):
    say(f"case j2 {etype} quadrature order 7")
    prm = dict(E=200.0, nu=0.3, sig0=0.25, H=1.0)
    m = _mesh(etype)
    C = m.elements()
    Q = fem.quadrature(etype, Q7)[0].numel()
    g = torch.Generator().manual_seed(11)
    ep = 1e-4 * torch.randn(C.shape[0], Q, 3, 3, generator=g, dtype=torch.float64)
    ep = 0.5 * (ep + ep.transpose(-1, -2))
    ep = ep - ep.diagonal(dim1=-2, dim2=-1).sum(-1)[..., None, None] / 3.0 * torch.eye(3, dtype=torch.float64)
    al = 1e-4 * torch.rand(C.shape[0], Q, generator=g, dtype=torch.float64)
    u = _smooth(m.points, 6e-4 if etype == "tetra" else 3e-4, 12)
    D = m.points.shape[1]
    e_ref = fem.j2_energy(m.points, C, etype, Q7, u, ep, al, **prm)
    res_e, tan_e = fem.j2_element_derivatives(m.points, C, etype, Q7, u, ep, al, **prm)
    res = torch.zeros(m.n_points, D, dtype=torch.float64).index_add(0, C.reshape(-1), res_e.reshape(-1, D))
    row, col, eids = fem.pattern(C.numpy(), m.n_points)
    tan = fem._index_add(tan_e, torch.from_numpy(eids), row.size)
    v, r, c = fem.block_coo(tan, torch.from_numpy(row), torch.from_numpy(col))
    model = tm.J2Plasticity.from_mesh(m.to(DEV), quadrature_order=Q7, **prm)
    hist = {"eps_p_n": {etype: ep.to(DEV)}, "alpha_n": {etype: al.to(DEV)}}
    energy = model.energy(point_data={"displacement": u.to(DEV)}, element_data=hist)
    assert abs(energy.item() - e_ref.item()) <= 1e-12 * abs(e_ref.item())
    R, K = model.residual_and_tangent(u.to(DEV), ep.to(DEV), al.to(DEV))
    ok, err = _within(R, res.reshape(-1), 1e-11, 1e-12)
    assert ok, f"residual {err:.2e}"
    assert torch.equal(K.row.cpu(), r) and torch.equal(K.col.cpu(), c)
    ok, err = _within(K.values, v, 1e-10, 1e-12)
    assert ok, f"tangent {err:.2e}"
    say(f"  energy {energy.item():.6e}  tangent max rel err {err:.2e}")
