"""Quadrilateral and hexahedral cells (quad, quad9, hexahedron, hexahedron27)
and cubic simplices (triangle10, tetra20) on the HIP kernels against the CPU
oracle.

Tolerances (float64): values within rtol 1e-12 plus a floor of 1e-14 max|ref|
for cancelling entries; Neo-Hookean / J2 / contact as in their own tests
(rtol 1e-11 / 1e-10, floor 1e-12 max|ref|).
"""
import math
import sys

import numpy as np
import pytest
import torch

import tensormesh_amd as tm
from oracle import fem
from tensormesh_amd.mesh import unit_cube_hex, unit_cube_tet, unit_square_quad, unit_square_tri

pytestmark = pytest.mark.gpu
DEV = "cuda"
TENSOR = ["quad", "quad9", "hexahedron", "hexahedron27"]
CUBIC = ["triangle10", "tetra20"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mesh(etype, n=None, seed=1):
    if etype in CUBIC:
        return (unit_square_tri(n or 3, 0.15, seed, order=3) if etype == "triangle10"
                else unit_cube_tet(n or 2, 0.1, seed, order=3))
    order = 2 if etype in ("quad9", "hexahedron27") else 1
    if etype.startswith("quad"):
        return unit_square_quad(n or 4, 0.2, seed, order=order)
    return unit_cube_hex(n or 3, 0.2, seed, order=order)


def _within(a, b, rtol, floor):
    a, b = a.detach().double().cpu(), b.detach().double().cpu()
    err = (a - b).abs()
    return bool((err <= rtol * b.abs() + floor * b.abs().max()).all()), err.max().item() / max(
        b.abs().max().item(), 1e-300)


def _assembled(Ke, C, n_points):
    row, col, eids = fem.pattern(C.numpy(), n_points)
    return fem._index_add(Ke.reshape((-1,) + tuple(Ke.shape[3:])), torch.from_numpy(eids), row.size), row, col


@pytest.mark.parametrize("etype,qo", [(e, q) for e in TENSOR for q in (2, 3)] + [(e, q) for e in CUBIC for q in (4, 6)])
def test_tensor_bilinear_forms_against_oracle(etype, qo):
    say(f"case tensor bilinear {etype} qorder {qo}")
    m = _mesh(etype)
    C = m.elements()
    md = m.to(DEV)
    kappa_e = torch.linspace(0.5, 2.0, C.shape[0], dtype=torch.float64)
    kappa_n = 1.0 + m.points[:, 0] ** 2
    cases = [("laplace", tm.LaplaceElementAssembler, {}, {}),
             ("diffusion_elem", tm.DiffusionElementAssembler, {"element_data": {"kappa": kappa_e.to(DEV)}},
              {"kappa": kappa_e}),
             ("diffusion_node", tm.DiffusionElementAssembler, {"point_data": {"kappa": kappa_n.to(DEV)}},
              {"kappa": kappa_n}),
             ("mass", tm.MassElementAssembler, {}, {})]
    for form, cls, call_kw, ref_kw in cases:
        K = cls.from_mesh(md, quadrature_order=qo)(**call_kw)
        ref, row, col = _assembled(fem.element_integrals(m.points, C, etype, qo, form, **ref_kw), C, m.n_points)
        assert np.array_equal(K.row.cpu().numpy(), row) and np.array_equal(K.col.cpu().numpy(), col)
        ok, err = _within(K.values, ref, 1e-12, 1e-14)
        assert ok, f"{form}: {err:.2e}"
    K = tm.LinearElasticityElementAssembler.from_mesh(md, quadrature_order=qo, E=2.0, nu=0.3)()
    Ke = fem.element_integrals(m.points, C, etype, qo, "linear_elasticity", E=2.0, nu=0.3)
    ref, row, col = _assembled(Ke, C, m.n_points)
    v, r, c = fem.block_coo(ref, torch.from_numpy(row), torch.from_numpy(col))
    assert torch.equal(K.row.cpu(), r) and torch.equal(K.col.cpu(), c)
    ok, err = _within(K.values, v, 1e-12, 1e-14)
    assert ok, f"elasticity: {err:.2e}"


@pytest.mark.parametrize("etype", TENSOR + CUBIC)
def test_tensor_loads_and_facet_loads_against_oracle(etype):
    say(f"case tensor loads {etype}")
    m = _mesh(etype)
    md = m.to(DEV)
    f = torch.sin(3 * m.points[:, 0]) + m.points[:, 1] ** 2

    class Source(tm.NodeAssembler):
        def forward(self, v, f):
            return f * v

    b = Source.from_mesh(md, quadrature_order=3)(point_data={"f": f.to(DEV)})
    want = fem.assemble_linear(m.points, m.elements(), etype, 3, f_nodal=f)
    ok, err = _within(b, want, 1e-12, 1e-14)
    assert ok, f"load: {err:.2e}"

    class Traction(tm.FacetAssembler):
        def forward(self, v, f):
            return f * v

    bf = Traction.from_mesh(md, quadrature_order=3)(point_data={"f": f.to(DEV)})
    wf = fem.assemble_facet_linear(m.points, m.elements(), m.boundary_mask, etype, 3, f_nodal=f)
    ok, err = _within(bf, wf, 1e-12, 1e-14)
    assert ok, f"facet load: {err:.2e}"


# tetra20 and hexahedron27 (60 and 81 element dofs) run the one-element-per-
# workgroup kernels; hexahedron27 at order 3 (27 points) and tetra20 at order 6
# (15 points) fill their larger quadrature arrays
@pytest.mark.parametrize("etype,qo", [("quad", 2), ("quad9", 2), ("hexahedron", 2), ("triangle10", 2),
                                      ("hexahedron27", 2), ("hexahedron27", 3), ("tetra20", 2), ("tetra20", 6)])
def test_tensor_neohookean_against_oracle(etype, qo):
    say(f"case tensor neo-hookean {etype} qorder {qo}")
    m = _mesh(etype, n=3 if etype.startswith(("quad", "tri")) else 2)
    P = m.points
    g = torch.Generator().manual_seed(5)
    u = 0.05 * torch.stack([torch.sin(P[:, 1]), P[:, 0] * P[:, -1], 0.5 * P[:, 0] ** 2][:P.shape[1]], 1)
    u = u + 0.002 * torch.randn(P.shape, generator=g, dtype=torch.float64)
    res_e, tan_e = fem.neohookean_element_tangent(P, m.elements(), etype, qo, u, E=1.0, nu=0.3)
    D = P.shape[1]
    res = torch.zeros(m.n_points, D, dtype=torch.float64).index_add(0, m.elements().reshape(-1),
                                                                     res_e.reshape(-1, D)).reshape(-1)
    tan, row, col = _assembled(tan_e, m.elements(), m.n_points)
    v, r, c = fem.block_coo(tan, torch.from_numpy(row), torch.from_numpy(col))
    model = tm.NeoHookeanModel.from_mesh(m.to(DEV), quadrature_order=qo, E=1.0, nu=0.3)
    R, K = model.residual_and_tangent(u.to(DEV))
    ok, err = _within(R, res, 1e-11, 1e-13)
    assert ok, f"residual: {err:.2e}"
    assert torch.equal(K.row.cpu(), r) and torch.equal(K.col.cpu(), c)
    ok, err = _within(K.values, v, 1e-11, 1e-13)
    assert ok, f"tangent: {err:.2e}"


# displacement amplitudes that leave about a third to two fifths of the points yielding
J2_AMP = {"hexahedron": 3e-4, "hexahedron27": 1e-4, "tetra20": 5e-5}


@pytest.mark.parametrize("etype,qo", [("hexahedron", 2), ("hexahedron27", 3), ("tetra20", 4)])
def test_tensor_j2_against_oracle(etype, qo):
    say(f"case tensor J2 {etype} qorder {qo}")
    prm = dict(E=200.0, nu=0.3, sig0=0.25, H=1.0)
    m = _mesh(etype, n=2, seed=3)
    C = m.elements()
    Q = fem.quadrature(etype, qo)[0].shape[0]
    g = torch.Generator().manual_seed(7)
    ep = 1e-4 * torch.randn(C.shape[0], Q, 3, 3, generator=g, dtype=torch.float64)
    ep = 0.5 * (ep + ep.transpose(-1, -2))
    al = 1e-4 * torch.rand(C.shape[0], Q, generator=g, dtype=torch.float64)
    u = J2_AMP[etype] * torch.randn(m.points.shape, generator=g, dtype=torch.float64)
    e_ref = fem.j2_energy(m.points, C, etype, qo, u, ep, al, **prm)
    res_e, tan_e = fem.j2_element_derivatives(m.points, C, etype, qo, u, ep, al, **prm)
    res = torch.zeros(m.n_points, 3, dtype=torch.float64).index_add(0, C.reshape(-1), res_e.reshape(-1, 3))
    tan, row, col = _assembled(tan_e, C, m.n_points)
    v, r, c = fem.block_coo(tan, torch.from_numpy(row), torch.from_numpy(col))
    model = tm.J2Plasticity.from_mesh(m.to(DEV), quadrature_order=qo, **prm)
    hist = {"eps_p_n": ep.to(DEV), "alpha_n": al.to(DEV)}
    e = model.energy(point_data={"displacement": u.to(DEV)}, element_data=hist)
    assert abs(e.item() - e_ref.item()) <= 1e-12 * abs(e_ref.item())
    R, K = model.residual_and_tangent(u.to(DEV), ep.to(DEV), al.to(DEV))
    ok, err = _within(R, res.reshape(-1), 1e-11, 1e-12)
    assert ok, f"residual: {err:.2e}"
    ok, err = _within(K.values, v, 1e-10, 1e-12)
    assert ok, f"tangent: {err:.2e}"


@pytest.mark.parametrize("etype", ["quad9", "hexahedron", "hexahedron27", "triangle10", "tetra20"])
def test_tensor_contact_against_oracle(etype):
    say(f"case tensor contact {etype}")
    m = _mesh(etype, n=3 if etype.startswith(("quad", "tri")) else 2)
    D = m.points.shape[1]
    n = [0.2, 1.0] if D == 2 else [0.1, 1.0, -0.2]
    s = math.sqrt(sum(x * x for x in n))
    n = [x / s for x in n]
    g = torch.Generator().manual_seed(9)
    x = m.points + 0.02 * torch.randn(m.points.shape, generator=g, dtype=torch.float64)
    x[:, 1] -= 0.2
    e_ref = fem.contact_energy(m.points, x, m.elements(), m.boundary_mask, etype, 2, 30.0, n, 0.05)
    res_f, _, elem, _ = fem.contact_facet_derivatives(m.points, x, m.elements(), m.boundary_mask, etype, 2, 30.0,
                                                      n, 0.05)
    dofs = m.elements()[torch.from_numpy(elem)]
    grad = torch.zeros(m.n_points, D, dtype=torch.float64).index_add(0, dofs.reshape(-1), res_f.reshape(-1, D))
    model = tm.PenaltyContact.from_mesh(m.to(DEV), quadrature_order=2, kappa=30.0, normal=n, offset=0.05)
    e = model.energy(points=x.to(DEV))
    assert e_ref.item() > 0 and abs(e.item() - e_ref.item()) <= 1e-12 * e_ref.item()
    ok, err = _within(model.residual(x.to(DEV)), grad.reshape(-1), 1e-11, 1e-13)
    assert ok, f"gradient: {err:.2e}"


def test_hexahedron27_elasticity_float32_against_float64():
    say("case hexahedron27 elasticity fp32 vs fp64")
    m = _mesh("hexahedron27", n=3)
    K64 = tm.LinearElasticityElementAssembler.from_mesh(m.to(DEV), quadrature_order=3, E=1.0, nu=0.3)()
    m32 = tm.Mesh(m.points.float(), {"hexahedron27": m.elements()}, m.boundary_mask).to(DEV)
    K32 = tm.LinearElasticityElementAssembler.from_mesh(m32, quadrature_order=3, E=1.0, nu=0.3)()
    d = K32.values.double() - K64.values
    normwise = (d.norm() / K64.values.norm()).item()
    say(f"  normwise {normwise:.2e}")
    assert normwise < 1e-6
