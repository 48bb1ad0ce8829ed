"""Penalty contact on the HIP kernel (tm_contact_penalty) against the CPU
oracle (oracle/fem.py, penalty-contact section).

Tolerances (float64): energy rtol 1e-12 (a sum of non-negative terms);
gradient |d| <= 1e-11 |ref| + 1e-13 max|ref|; Hessian |d| <= 1e-11 |ref| +
1e-13 max|ref| (the kernel's closed form against autograd).
"""
import math
import sys

import pytest
import torch

import tensormesh_amd as tm
from oracle import fem
from tensormesh_amd.mesh import Mesh, quadratic_arrays, unit_cube_tet_arrays, unit_square_tri_arrays

pytestmark = pytest.mark.gpu
DEV = "cuda"


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mesh(etype, n=3, seed=1):
    if etype.startswith("tri"):
        p, c, bm = unit_square_tri_arrays(n, 0.15, seed)
    else:
        p, c, bm = unit_cube_tet_arrays(n, 0.1, seed)
    if etype in ("triangle6", "tetra10"):
        p, c, bm = quadratic_arrays(p, c, bm)
    return Mesh(torch.from_numpy(p), {etype: torch.from_numpy(c)}, torch.from_numpy(bm))


def _within(a, b, rtol, floor):
    a, b = a.detach().double().cpu(), b.detach().double().cpu()
    err = (a - b).abs()
    return bool((err <= rtol * b.abs() + floor * b.abs().max()).all()), err.max().item() / max(
        b.abs().max().item(), 1e-300)


def _plane(D):
    n = [0.25, 1.0] if D == 2 else [0.2, 1.0, -0.15]
    s = math.sqrt(sum(v * v for v in n))
    return [v / s for v in n], 0.02


def _positions(m, seed):
    g = torch.Generator().manual_seed(seed)
    x = m.points + 0.02 * torch.randn(m.points.shape, generator=g, dtype=torch.float64)
    x[:, 1] -= 0.15
    return x


@pytest.mark.parametrize("etype", ["triangle", "triangle6", "tetra", "tetra10"])
def test_contact_energy_gradient_hessian_against_oracle(etype):
    say(f"case penalty contact {etype}")
    m = _mesh(etype)
    D = m.points.shape[1]
    nrm, off = _plane(D)
    x = _positions(m, 3)
    C, BM = m.elements(), m.boundary_mask
    kappa = 40.0
    e_ref = fem.contact_energy(m.points, x, C, BM, etype, 2, kappa, nrm, off)
    res_f, hes_f, elem, fac = fem.contact_facet_derivatives(m.points, x, C, BM, etype, 2, kappa, nrm, off)
    dofs = C[torch.from_numpy(elem)]
    grad = torch.zeros(m.n_points, D, dtype=torch.float64).index_add(0, dofs.reshape(-1), res_f.reshape(-1, D))
    # Hessian: facet blocks onto the cell mesh's pattern entries
    row, col, eids = fem.pattern(C.numpy(), m.n_points)
    B = C.shape[1]
    ar = torch.arange(B)
    eslot = (torch.from_numpy(elem)[:, None, None] * B * B + ar[:, None] * B + ar[None, :]).reshape(-1)
    tan = torch.zeros(row.size, D, D, dtype=torch.float64).index_add(
        0, torch.from_numpy(eids)[eslot], hes_f.reshape(-1, D, D))
    v, r, c = fem.block_coo(tan, torch.from_numpy(row), torch.from_numpy(col))

    model = tm.PenaltyContact.from_mesh(m.to(DEV), quadrature_order=2, kappa=kappa, normal=nrm, offset=off)
    xd = x.to(DEV).requires_grad_(True)
    energy = model.energy(points=xd)
    assert energy.device.type == "cuda"
    assert e_ref.item() > 0 and abs(energy.item() - e_ref.item()) <= 1e-12 * e_ref.item()
    energy.backward()
    ok, err = _within(xd.grad.reshape(-1), grad.reshape(-1), 1e-11, 1e-13)
    assert ok, f"energy backward: {err:.2e}"
    R, K = model.residual_and_tangent(x.to(DEV))
    ok, err = _within(R, grad.reshape(-1), 1e-11, 1e-13)
    assert ok, f"gradient: {err:.2e}"
    assert torch.equal(K.row.cpu(), r) and torch.equal(K.col.cpu(), c)
    ok, err = _within(K.values, v, 1e-11, 1e-13)
    assert ok, f"hessian: {err:.2e}"
    say(f"  energy {energy.item():.6e}  facets {len(elem)}  hessian err {err:.2e}")


def test_contact_clear_of_the_plane_is_zero_and_layouts_add():
    """Lifted clear of the plane nothing is in contact: zero energy, gradient
    and Hessian values; the Hessian shares the element tangent's layout, so
    the two add."""
    say("case penalty contact clear / layout")
    m = _mesh("tetra", n=3).to(DEV)
    model = tm.PenaltyContact.from_mesh(m, kappa=10.0, obstacle_y=-1.0)
    x = m.points.clone()
    e = model.energy(points=x)
    R, K = model.residual_and_tangent(x)
    assert e.item() == 0.0 and R.abs().max().item() == 0.0 and K.values.abs().max().item() == 0.0
    lin = tm.LinearElasticityElementAssembler.from_mesh(m, E=1.0, nu=0.3)()
    assert K.has_same_layout(lin)
    S = lin + K
    assert torch.equal(S.values, lin.values)


def test_contact_float32_against_float64():
    say("case penalty contact tetra10 fp32 vs fp64")
    m = _mesh("tetra10", n=3)
    D = 3
    nrm, off = _plane(D)
    x = _positions(m, 5)
    m64 = tm.PenaltyContact.from_mesh(m.to(DEV), kappa=40.0, normal=nrm, offset=off)
    m32m = Mesh(m.points.float(), {"tetra10": m.elements()}, m.boundary_mask).to(DEV)
    m32 = tm.PenaltyContact.from_mesh(m32m, kappa=40.0, normal=nrm, offset=off)
    R64, K64 = m64.residual_and_tangent(x.to(DEV))
    R32, K32 = m32.residual_and_tangent(x.float().to(DEV))
    e64, e32 = m64.energy(points=x.to(DEV)).item(), m32.energy(points=x.float().to(DEV)).item()
    dk = ((K32.values.double() - K64.values).norm() / K64.values.norm()).item()
    dr = ((R32.double() - R64).norm() / R64.norm()).item()
    say(f"  fp32 vs fp64: hessian {dk:.2e} gradient {dr:.2e} energy {abs(e32 - e64) / e64:.2e}")
    # penetrations are differences of positions ~ 0.1: float32 rounding of x
    # gives ~1e-6 relative error in them
    assert dk < 1e-5 and dr < 1e-4 and abs(e32 - e64) < 1e-4 * e64


def test_no_selected_facet_through_the_abi():
    """tm_contact_penalty with n_sel = 0: a full Hessian view is zeroed, a
    subset view (entry set) leaves the output to the caller's zeroing."""
    import ctypes
    from tensormesh_amd import _native as nat
    say("case contact n_sel=0 through the C-ABI")
    L = nat.lib()
    dev = torch.device(DEV)
    m = _mesh("tetra", n=2).to(DEV)
    conn = m.elements().to(torch.int32).contiguous()
    mv = nat.MeshView(nat.CELL["tetra"], nat.TM_F64, conn.shape[0], m.points.shape[0], conn.data_ptr(),
                      m.points.data_ptr())
    fv = nat.FacetView(0, None, None)
    tab = torch.zeros(64, dtype=torch.float64, device=DEV)
    plane = nat.ContactPlane(1.0, (ctypes.c_double * 3)(0.0, 1.0, 0.0), 0.0)
    nnz, D = 12, 3
    idx = torch.arange(nnz, dtype=torch.int32, device=DEV)
    seg_ptr = torch.zeros(nnz + 1, dtype=torch.int32, device=DEV)
    tslot = torch.zeros(9, dtype=torch.float64, device=DEV)
    hess = torch.full((2 * nnz * D * D,), 7.0, dtype=torch.float64, device=DEV)
    st = nat.stream_of(dev)
    sub = nat.PatternView(nnz, None, None, seg_ptr.data_ptr(), idx.data_ptr(), 0, None, idx.data_ptr(),
                          idx.data_ptr())
    nat.check(L.tm_contact_penalty(ctypes.byref(mv), ctypes.byref(fv), tab.data_ptr(), 1, ctypes.byref(plane),
                                   m.points.data_ptr(), None, None, None, None, None, ctypes.byref(sub),
                                   tslot.data_ptr(), hess.data_ptr(), st), "tm_contact_penalty")
    torch.cuda.synchronize()
    assert bool((hess == 7.0).all()), "a subset view must not write the Hessian"
    full = nat.PatternView(nnz, None, None, seg_ptr.data_ptr(), idx.data_ptr(), 0, None, None, None)
    nat.check(L.tm_contact_penalty(ctypes.byref(mv), ctypes.byref(fv), tab.data_ptr(), 1, ctypes.byref(plane),
                                   m.points.data_ptr(), None, None, None, None, None, ctypes.byref(full),
                                   tslot.data_ptr(), hess.data_ptr(), st), "tm_contact_penalty")
    torch.cuda.synchronize()
    assert bool((hess[:nnz * D * D] == 0.0).all()) and bool((hess[nnz * D * D:] == 7.0).all())
