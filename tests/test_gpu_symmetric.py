"""Symmetric element blocks (TM_FLAG_UPPER) against full blocks.

With ``symmetric_slots`` (the default) the element-matrix, Neo-Hookean and J2
kernels store only the blocks (i, j) with i <= j, and the segment reduction
reads every other block transposed through the pattern's symmetric segment
list (tm_symmetric_segments).  The kernels form their products so that block
(j, i) is the exact transpose of block (i, j), and every entry still sums its
contributions in ascending element order, so the assembled values must equal
those of the full blocks (``symmetric_slots = False``, as the reference
stores them) bit for bit.  Agreement of the full blocks with the oracle is
covered by the parity tests of each form.
"""
import sys

import pytest
import torch

import tensormesh_amd as tm
from oracle import fem
from tensormesh_amd.mesh import unit_cube_hex, unit_cube_tet, unit_square_quad, unit_square_tri

pytestmark = pytest.mark.gpu
DEV = "cuda"


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mesh(etype, seed=1):
    if etype in ("triangle", "triangle6", "triangle10"):
        return unit_square_tri(4, 0.15, seed, order={"triangle": 1, "triangle6": 2, "triangle10": 3}[etype])
    if etype in ("tetra", "tetra10", "tetra20"):
        return unit_cube_tet(2, 0.1, seed, order={"tetra": 1, "tetra10": 2, "tetra20": 3}[etype])
    if etype in ("quad", "quad9"):
        return unit_square_quad(4, 0.2, seed, order=2 if etype == "quad9" else 1)
    return unit_cube_hex(2, 0.2, seed, order=2 if etype == "hexahedron27" else 1)


def _both(make, run):
    """run(make()) with symmetric and with full element blocks."""
    out = []
    for sym in (True, False):
        A = make()
        A.symmetric_slots = sym
        out.append(run(A))
    return out


def _same(a, b, what):
    a, b = a.detach(), b.detach()
    assert a.shape == b.shape, what
    assert torch.equal(a, b), f"{what}: max |diff| {(a - b).abs().max().item():.3e}"


@pytest.mark.parametrize("etype", ["triangle6", "tetra10", "triangle10", "tetra20", "quad", "quad9", "hexahedron",
                                   "hexahedron27", "tetra"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_symmetric_blocks_bilinear_forms(etype, dtype):
    say(f"case symmetric bilinear {etype} {dtype}")
    m = _mesh(etype)
    md = tm.Mesh(m.points.to(dtype), {etype: m.elements()}, m.boundary_mask).to(DEV)
    E = m.elements().shape[0]
    kappa = torch.linspace(0.5, 2.0, E, dtype=dtype, device=DEV)
    qo = 3 if etype in ("triangle10", "tetra20") else 2
    cases = [("linear_elasticity", lambda: tm.LinearElasticityElementAssembler.from_mesh(md, quadrature_order=qo,
                                                                                         E=2.0, nu=0.3), {})]
    if etype != "tetra":   # P1 scalar forms run the P1 kernels, which store every block
        cases += [("laplace", lambda: tm.LaplaceElementAssembler.from_mesh(md, quadrature_order=qo), {}),
                  ("diffusion", lambda: tm.DiffusionElementAssembler.from_mesh(md, quadrature_order=qo),
                   {"element_data": {"kappa": kappa}}),
                  ("mass", lambda: tm.MassElementAssembler.from_mesh(md, quadrature_order=qo), {})]
    for name, make, kw in cases:
        Ks, Kf = _both(make, lambda A: A(**kw))
        assert torch.equal(Ks.row, Kf.row) and torch.equal(Ks.col, Kf.col)
        _same(Ks.values, Kf.values, f"{etype} {name}")


@pytest.mark.parametrize("etype,qo", [("tetra10", 2), ("triangle6", 2), ("hexahedron", 2), ("tetra20", 2),
                                      ("hexahedron27", 3)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_symmetric_blocks_neohookean(etype, qo, dtype):
    say(f"case symmetric neo-hookean {etype} qorder {qo} {dtype}")
    m = _mesh(etype)
    P = m.points
    u = 0.05 * torch.stack([torch.sin(P[:, 1]), P[:, 0] * P[:, -1], 0.5 * P[:, 0] ** 2][:P.shape[1]], 1)
    md = tm.Mesh(P.to(dtype), {etype: m.elements()}, m.boundary_mask).to(DEV)
    ud = u.to(dtype).to(DEV)
    (Rs, Ks), (Rf, Kf) = _both(lambda: tm.NeoHookeanModel.from_mesh(md, quadrature_order=qo, E=1.0, nu=0.3),
                               lambda A: A.residual_and_tangent(ud))
    _same(Rs, Rf, "residual")
    _same(Ks.values, Kf.values, "tangent")


@pytest.mark.parametrize("etype,qo", [("tetra", 2), ("triangle", 2), ("tetra10", 2), ("hexahedron27", 3)])
def test_symmetric_blocks_j2(etype, qo):
    say(f"case symmetric J2 {etype} qorder {qo}")
    m = _mesh(etype, seed=3)
    E = m.elements().shape[0]
    Q = fem.quadrature(etype, qo)[0].shape[0]
    g = torch.Generator().manual_seed(7)
    ep = 1e-4 * torch.randn(E, Q, 3, 3, generator=g, dtype=torch.float64)
    ep = (0.5 * (ep + ep.transpose(-1, -2))).to(DEV)
    al = (1e-4 * torch.rand(E, Q, generator=g, dtype=torch.float64)).to(DEV)
    u = (2e-4 * torch.randn(m.points.shape, generator=g, dtype=torch.float64)).to(DEV)
    (Rs, Ks), (Rf, Kf) = _both(lambda: tm.J2Plasticity.from_mesh(m.to(DEV), quadrature_order=qo, E=200.0, nu=0.3,
                                                                  sig0=0.25, H=1.0),
                               lambda A: A.residual_and_tangent(u, ep, al))
    _same(Rs, Rf, "residual")
    _same(Ks.values, Kf.values, "tangent")


def test_symmetric_segment_list():
    """tm_symmetric_segments on a tetra10 pattern: every slot becomes the packed
    upper block of (min(i, j), max(i, j)), with bit 31 set when i > j."""
    m = _mesh("tetra10").to(DEV)
    A = tm.LaplaceElementAssembler.from_mesh(m, quadrature_order=2)
    A()
    pat = A._pattern_on(torch.device(DEV))
    seg = pat.seg_slot.long().cpu()
    sym = pat.sym_slot().long().cpu()
    B = pat.n_basis
    e, r = seg // (B * B), seg % (B * B)
    i, j = r // B, r % B
    a, b = torch.minimum(i, j), torch.maximum(i, j)
    packed = e * (B * (B + 1) // 2) + a * B - a * (a - 1) // 2 + (b - a)
    want = torch.where(i <= j, packed, packed - 2 ** 31)
    assert torch.equal(sym, want)
    assert int((i > j).sum()) > 0 and int((i <= j).sum()) > 0


def test_symmetric_flag_is_rejected_where_blocks_are_stored_whole():
    """TM_FLAG_UPPER on the P1 scalar kernels, the MFMA tangent and the
    block-row path fails with TM_ERR_ARG before any launch (those store every
    element block, which a symmetric segment list would misread)."""
    import ctypes
    from tensormesh_amd import _native as nat
    from tensormesh_amd.assemble import _n_quad
    L = nat.lib()
    m = _mesh("tetra").to(DEV)
    A = tm.LaplaceElementAssembler.from_mesh(m, quadrature_order=1)
    tab, tab_dev = A._tables(torch.device(DEV))
    mv = A._mesh_view(m.points)
    E, B = A._conn32(torch.device(DEV)).shape
    form = nat.FormDesc(nat.FORM["laplace"], _n_quad("tetra", 1), tab.ctypes.data, tab_dev.data_ptr(), None, 0.0,
                        0.0, nat.TM_FLAG_UPPER)
    slot = torch.empty(E * B * B, dtype=torch.float64, device=DEV)
    assert L.tm_element_matrices(ctypes.byref(mv), ctypes.byref(form), slot.data_ptr(),
                                 nat.stream_of(torch.device(DEV))) == -1
    assert "P1" in L.tm_last_error().decode()
    m10 = _mesh("tetra10")
    md = tm.Mesh(m10.points.float(), {"tetra10": m10.elements()}, m10.boundary_mask).to(DEV)
    model = tm.NeoHookeanModel.from_mesh(md, quadrature_order=2, E=1.0, nu=0.3)
    tab, tab_dev = model._tables(torch.device(DEV))
    mv = model._mesh_view(md.points)
    E, B = model._conn32(torch.device(DEV)).shape
    u = torch.zeros_like(md.points)
    form = nat.FormDesc(0, _n_quad("tetra10", 2), tab.ctypes.data, tab_dev.data_ptr(), None, 1.0, 0.3,
                        nat.TM_FLAG_UPPER | nat.TM_FLAG_MFMA)
    pat = model._pattern_on(torch.device(DEV))
    pv = pat.view(torch.float32, 9, symmetric=True)
    tslot = torch.empty(E * B * B * 9, dtype=torch.float32, device=DEV)
    tan = torch.empty(pat.nnz * 9, dtype=torch.float32, device=DEV)
    st = L.tm_assemble_neohookean(ctypes.byref(mv), ctypes.byref(form), u.data_ptr(), ctypes.byref(pv), None,
                                  tslot.data_ptr(), tan.data_ptr(), None, None, nat.stream_of(torch.device(DEV)))
    assert st == -1 and "TM_FLAG_UPPER" in L.tm_last_error().decode()


@pytest.mark.parametrize("etype", ["tetra10", "hexahedron", "triangle6", "tetra20", "tetra"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_contribution_ordered_slots_equal_gathered_slots(etype, dtype):
    """Symmetric blocks written in contribution order (slot_dest, contiguous
    segment sums) against the same blocks gathered through the segment list:
    bit-identical linear-elasticity, Neo-Hookean and J2 tangents."""
    from tensormesh_amd.pattern import Pattern
    say(f"case contiguous slots {etype} {dtype}")
    m = _mesh(etype, seed=3)
    md = tm.Mesh(m.points.to(dtype), m.cells, m.boundary_mask).to(DEV)
    D = md.points.shape[1]
    P = md.points
    u = 0.03 * torch.stack([torch.sin(P[:, 1]), P[:, 0] * P[:, -1], 0.5 * P[:, 0] ** 2][:D], 1)
    try:
        res = {}
        for contiguous in (True, False):
            Pattern.contiguous_slots = contiguous
            L = tm.LinearElasticityElementAssembler.from_mesh(md, quadrature_order=2, E=1.0, nu=0.3)
            L.symmetric_slots = True
            nh = tm.NeoHookeanModel.from_mesh(md, quadrature_order=2, E=1.0, nu=0.3)
            nh.symmetric_slots = True
            j2 = tm.J2Plasticity.from_mesh(md, quadrature_order=2, E=200.0, nu=0.3, sig0=0.25, H=1.0)
            j2.symmetric_slots = True
            res[contiguous] = (L().values, nh.tangent(u), j2.tangent(0.1 * u))
            if contiguous:
                assert L._pattern_on(torch.device(DEV)).slot_dest() is not None
    finally:
        Pattern.contiguous_slots = True
    for a, b, what in zip(res[True], res[False], ("elasticity", "neo-hookean", "j2")):
        a = a.values if isinstance(a, tm.SparseMatrix) else a
        b = b.values if isinstance(b, tm.SparseMatrix) else b
        _same(a, b, what)
