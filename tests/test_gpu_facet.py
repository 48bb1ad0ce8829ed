"""Boundary-facet loads (FacetAssembler) on the HIP kernel against the oracle."""
import pytest
import torch

import tensormesh_amd as tm
from oracle import fem
from tensormesh_amd.mesh import quadratic_arrays, unit_cube_tet_arrays, unit_square_tri_arrays

pytestmark = pytest.mark.gpu
DEV = "cuda"


def say(msg):
    print(msg, flush=True)


def _mesh(kind, order, mask):
    if kind == "tri":
        p, c, bm = unit_square_tri_arrays(16, 0.2, 5)
        et = "triangle" if order == 1 else "triangle6"
    else:
        p, c, bm = unit_cube_tet_arrays(5, 0.1, 5)
        et = "tetra" if order == 1 else "tetra10"
    if order == 2:
        p, c, bm = quadratic_arrays(p, c, bm)
    if mask == "face":
        bm = p[:, 0] <= 1e-12
    return torch.from_numpy(p), torch.from_numpy(c), torch.from_numpy(bm), et


class ScalarLoad(tm.FacetAssembler):
    def forward(self, v, f):
        return f * v


class VectorLoad(tm.FacetAssembler):
    def forward(self, v, g):
        return v[:, None] * g[None, :]


@pytest.mark.parametrize("kind,order", [("tri", 1), ("tri", 2), ("tet", 1), ("tet", 2)])
@pytest.mark.parametrize("mask", ["all", "face"])
def test_native_facet_loads_match_oracle(kind, order, mask, monkeypatch):
    say(f"case facet {kind} P{order} mask={mask}")
    P, C, BM, et = _mesh(kind, order, mask)
    D = P.shape[1]
    m = tm.Mesh(P, {et: C}, BM).to(DEV)
    # the built-in integrands must not fall back to the torch path
    monkeypatch.setattr(tm.FacetAssembler, "_torch", lambda *a, **k: (_ for _ in ()).throw(AssertionError("torch")))
    A = ScalarLoad.from_mesh(m, quadrature_order=2)
    f = torch.sin(3 * P[:, 0]) + P[:, 1] ** 2
    b = A(point_data={"f": f.to(DEV)})
    ref = fem.assemble_facet_linear(P, C, BM, et, 2, f_nodal=f)
    assert torch.allclose(b.cpu(), ref, rtol=1e-12, atol=1e-15)
    g = torch.stack([P[:, 0] * P[:, 1], torch.cos(P[:, 1]), P[:, 0] - 0.5][:D], 1)
    bg = VectorLoad.from_assembler(A)(point_data={"g": g.to(DEV)})
    refg = fem.assemble_facet_linear(P, C, BM, et, 2, f_nodal=g)
    assert torch.allclose(bg.cpu(), refg, rtol=1e-12, atol=1e-15)

    t = torch.tensor([0.25, -1.0, 0.5][:D], dtype=torch.float64)

    class Traction(tm.FacetAssembler):
        def forward(self, v):
            return v[:, None] * t.to(v.device, v.dtype)[None, :]

    bt = Traction.from_assembler(A)()
    reft = fem.assemble_facet_linear(P, C, BM, et, 2, const=t)
    assert torch.allclose(bt.cpu(), reft, rtol=1e-12, atol=1e-15)


def test_facet_float32_and_quadrature_orders():
    say("case facet float32 / orders")
    P, C, BM, et = _mesh("tet", 2, "all")
    f = torch.sin(3 * P[:, 0]) + P[:, 1] ** 2
    for q in (1, 2, 3):
        A = ScalarLoad.from_mesh(tm.Mesh(P, {et: C}, BM).to(DEV), quadrature_order=q)
        ref = fem.assemble_facet_linear(P, C, BM, et, q, f_nodal=f)
        b = A(point_data={"f": f.to(DEV)})
        assert torch.allclose(b.cpu(), ref, rtol=1e-12, atol=1e-15)
    m32 = tm.Mesh(P.float(), {et: C}, BM).to(DEV)
    b32 = ScalarLoad.from_mesh(m32, quadrature_order=2)(point_data={"f": f.float().to(DEV)})
    ref = fem.assemble_facet_linear(P.float().double(), C, BM, et, 2, f_nodal=f.float().double())
    assert b32.dtype == torch.float32
    assert (b32.double().cpu() - ref).abs().max() <= 1e-6 * ref.abs().max()


def test_facet_custom_integrand_on_device_torch_path():
    say("case facet custom integrand")
    P, C, BM, et = _mesh("tri", 2, "all")

    class Custom(tm.FacetAssembler):
        def forward(self, v, gradv, x):
            return (x * x).sum() * v + gradv[:, 1]

    b = Custom.from_mesh(tm.Mesh(P, {et: C}, BM).to(DEV), quadrature_order=2)()
    elem, fac = fem.facet_select(C.numpy(), BM.numpy(), et)
    phi, grad, fxw = fem.facet_geometry(P, C, elem, fac, et, 2)
    dofs = C[torch.from_numpy(elem)]
    xq = torch.einsum("fbd,fqb->fqd", P[dofs], phi)
    integrand = (xq * xq).sum(-1)[..., None] * phi + grad[..., 1]
    ref = torch.zeros(P.shape[0], dtype=torch.float64).index_add(
        0, dofs.reshape(-1), torch.einsum("fqb,fq->fb", integrand, fxw).reshape(-1))
    assert torch.allclose(b.cpu(), ref, rtol=1e-12, atol=1e-14)
