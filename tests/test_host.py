"""CPU tests of the host logic: the C-ABI library, the host pattern builder,
form recognition, the generic torch form path and the API containers."""
import ctypes
import math
import os
import re
from pathlib import Path

import numpy as np
import pytest
import torch

import tensormesh_amd as tm
from oracle import fem
from tensormesh_amd import _native as nat
from tensormesh_amd.forms import match_bilinear, match_linear
from tensormesh_amd.mesh import lognormal_coefficient, quadratic_arrays, unit_cube_tet, unit_cube_tet_arrays, unit_square_tri
from tensormesh_amd.pattern import Pattern, layout_hash

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_every_header_symbol():
    L = nat.lib()
    header = (ROOT / "include" / "tensormesh_amd.h").read_text()
    names = set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(tm_\w+)\s*\(", header, flags=re.M))
    assert names, "no declarations parsed"
    assert names == set(nat.SIGNATURES), names ^ set(nat.SIGNATURES)
    for n in names:
        assert hasattr(L, n)
    assert L.tm_abi_version() == nat.ABI_VERSION


def test_noalloc_load_rejects_sizes_beyond_its_array():
    """tm_assemble_linear_noalloc refuses slot values beyond its fixed device
    array with TM_ERR_SIZE before touching the device (host arithmetic only)."""
    L = nat.lib()
    B = 4
    n_elem = nat.LINEAR_SCRATCH_BYTES // (8 * B) + 1           # one float64 tetrahedron too many
    dummy = ctypes.c_int32(0)
    mesh = nat.MeshView(nat.CELL["tetra"], nat.TM_F64, n_elem, 8, ctypes.addressof(dummy), ctypes.addressof(dummy))
    tab = ctypes.c_double(0.0)
    form = nat.FormDesc(0, 1, ctypes.addressof(tab), ctypes.addressof(tab), None, 0.0, 0.0)
    iv = nat.IncidenceView(ctypes.addressof(dummy), ctypes.addressof(dummy), 0, None)
    c = ctypes.c_double(1.0)
    st = L.tm_assemble_linear_noalloc(ctypes.byref(mesh), ctypes.byref(form), ctypes.byref(c), ctypes.byref(iv),
                                      ctypes.addressof(dummy), None)
    assert st == -2 and b"scratch" in L.tm_last_error()
    mesh.dtype = nat.TM_F32                                      # the same count fits in float32
    assert nat.LINEAR_SCRATCH_BYTES // (4 * B) >= n_elem


def test_errors_cross_the_abi_as_status():
    L = nat.lib()
    conn = np.array([[0, 1, 5]], dtype=np.int32)
    nnz = ctypes.c_int64(0)
    st = L.tm_pattern_nnz(conn.ctypes.data, 1, 3, 3, ctypes.byref(nnz))
    assert st == -1 and b"outside" in L.tm_last_error()
    with pytest.raises(nat.NativeError):
        nat.check(st, "tm_pattern_nnz")


@pytest.mark.parametrize("kind,n", [("tri", 7), ("tet", 3), ("tet10", 2)])
def test_pattern_builder_matches_reference_construction(kind, n):
    if kind == "tri":
        m = unit_square_tri(n, 0.2, 1)
    elif kind == "tet":
        m = unit_cube_tet(n, 0.15, 1)
    else:
        m = unit_cube_tet(n, 0.0, 1, order=2)
    conn = m.elements()
    pat = Pattern.build(conn, m.n_points)
    row, col, eids = fem.pattern(conn.numpy(), m.n_points)
    assert np.array_equal(pat.row.numpy(), row) and np.array_equal(pat.col.numpy(), col)
    assert np.array_equal(pat.slot2nnz.numpy(), eids)
    # segments: ascending slots, exactly the slots mapped to each entry
    seg_ptr, seg_slot = pat.seg_ptr.numpy(), pat.seg_slot.numpy()
    assert seg_ptr[-1] == eids.size
    for k in range(0, pat.nnz, max(1, pat.nnz // 200)):
        s = seg_slot[seg_ptr[k]:seg_ptr[k + 1]]
        assert np.all(np.diff(s) > 0)
        assert np.array_equal(np.sort(np.nonzero(eids == k)[0]), s)
    edges = fem.reference_edges(conn, m.n_points)
    assert pat.hash == fem.layout_hash(edges[0], edges[1])


def test_incidence_is_index_add_order():
    m = unit_cube_tet(2, 0.1, 0)
    nd = tm.NodeAssembler.from_mesh(m)
    inc_ptr, inc = (t.numpy() for t in nd._incidence)
    flat = m.elements().numpy().reshape(-1)
    for n in range(m.n_points):
        assert np.array_equal(inc[inc_ptr[n]:inc_ptr[n + 1]], np.nonzero(flat == n)[0])


def test_form_recognition():
    class Lap(tm.ElementAssembler):
        def forward(self, gradu, gradv):
            return gradu @ gradv

    class Scaled(tm.ElementAssembler):
        def forward(self, gradu, gradv):
            return 2.5 * (gradu @ gradv)

    class Weird(tm.ElementAssembler):
        def forward(self, gradu, gradv, k):
            return k * gradu[0] * gradv[0]

    E = torch.ones(4)
    assert match_bilinear(Lap.forward.__get__(object()), 2, {}, {}, {}).form == "laplace"
    assert match_bilinear(Scaled.forward.__get__(object()), 2, {}, {}, {}) is None
    assert match_bilinear(Weird.forward.__get__(object()), 2, {"k": E}, {}, {}) is None
    f = lambda gradu, gradv, kappa: kappa * (gradu @ gradv)
    assert match_bilinear(f, 3, {"kappa": E}, {}, {}).form == "diffusion_elem"
    assert match_bilinear(f, 3, {}, {"kappa": torch.ones(5)}, {}).form == "diffusion_node"
    assert match_bilinear(lambda u, v: u * v, 3, {}, {}, {}).form == "mass"
    el = tm.LinearElasticityElementAssembler.__new__(tm.LinearElasticityElementAssembler)
    el.E, el.nu = 2.0, 0.25
    m = match_bilinear(el.forward, 3, {}, {}, {}, owner=el)
    assert m.form == "linear_elasticity" and (m.young, m.poisson) == (2.0, 0.25)
    assert match_linear(lambda v, f: f * v, {"f": torch.ones(3)}, {}).key == "f"
    lm = match_linear(tm.const_node_assembler(3.0)().forward if False else (lambda v: 3.0 * v), {}, {})
    assert lm.key is None and abs(lm.const - 3.0) < 1e-15
    assert match_linear(lambda x, v: x[0] * v, {"x": torch.ones(3, 2)}, {}) is None


def test_builtin_forms_on_cpu_tensors_run_the_torch_path():
    """On CPU tensors the built-in classes run the reference's generic torch
    path (as the reference does on the CPU), with the reference's values,
    pattern and layout digest; the HIP kernels serve device tensors only."""
    m = unit_square_tri(20, 0.2, 0)
    K = tm.LaplaceElementAssembler.from_mesh(m)()
    v, r, c, _ = fem.assemble_bilinear(m.points, m.elements(), "triangle", 2, "laplace")
    assert K.device.type == "cpu" and torch.equal(K.row, r) and torch.equal(K.col, c)
    assert torch.allclose(K.values, v, rtol=1e-13, atol=1e-15)
    edges = fem.reference_edges(m.elements(), m.n_points)
    assert K.layout_hash == fem.layout_hash(edges[0], edges[1])
    kap = torch.from_numpy(lognormal_coefficient(m.elements().shape[0], 5))
    Kd = tm.DiffusionElementAssembler.from_mesh(m)(element_data={"kappa": kap})
    vd = fem.assemble_bilinear(m.points, m.elements(), "triangle", 2, "diffusion_elem", kappa=kap)[0]
    assert torch.allclose(Kd.values, vd, rtol=1e-13, atol=1e-15)
    f = 2 * math.pi ** 2 * torch.sin(math.pi * m.points[:, 0]) * torch.sin(math.pi * m.points[:, 1])
    b = tm.SourceNodeAssembler.from_mesh(m)(point_data={"f": f})
    assert torch.allclose(b, fem.assemble_linear(m.points, m.elements(), "triangle", 2, f), rtol=1e-13, atol=1e-16)


def test_near_miss_forms_are_not_recognised():
    """Integrands that agree with a built-in at (almost) every point but are
    not polynomials keep the generic path: the polynomial guard rejects the
    operation, whatever the values at the probe points."""
    near = lambda gradu, gradv: torch.where(gradu[0] > 0.999999, 2.0 * (gradu @ gradv), gradu @ gradv)
    assert match_bilinear(near, 3, {}, {}, {}) is None

    def branch(gradu, gradv):
        if gradu[0] > 0.999999:
            return 2.0 * (gradu @ gradv)
        return gradu @ gradv
    assert match_bilinear(branch, 3, {}, {}, {}) is None
    clamp = lambda u, v: torch.clamp(u, -10.0, 10.0) * v          # equal to u * v on the probe range
    assert match_bilinear(clamp, 3, {}, {}, {}) is None
    detached = lambda gradu, gradv: gradu.detach() @ gradv
    assert match_bilinear(detached, 3, {}, {}, {}) is None
    via_abs = lambda v: (v.abs() + v - v.abs()) * 1.5             # = 1.5 v, through a non-polynomial op
    assert match_linear(via_abs, {}, {}) is None
    # polynomial spellings of the built-ins are still recognised
    assert match_bilinear(lambda gradu, gradv: (gradu * gradv).sum(), 3, {}, {}, {}).form == "laplace"
    assert match_bilinear(lambda gradu, gradv: torch.einsum("i,i->", gradu, gradv) / 1.0, 2, {}, {}, {}).form \
        == "laplace"


def test_recognition_cache_follows_closures_attributes_and_globals():
    """A cached recognition is redone when anything the integrand reads
    changes (closure cell, module attribute, tensor attribute in place)."""
    m = unit_square_tri(3, 0.0, 0)
    node = tm.NodeAssembler.from_mesh(m)
    scale = [1.0]
    fn = lambda v: scale[0] * v
    pd = {}
    assert abs(node._match(fn, pd, {}).const - 1.0) < 1e-15
    scale[0] = 5.0
    assert abs(node._match(fn, pd, {}).const - 5.0) < 1e-15

    def make(c):
        return lambda v: c * v
    f2 = make(2.0)
    assert abs(node._match(f2, pd, {}).const - 2.0) < 1e-15

    class Src(tm.NodeAssembler):
        def forward(self, v):
            return self.k * v
    src = Src.from_mesh(m)
    src.k = torch.tensor(2.0, dtype=torch.float64)
    assert abs(src._match(src.forward, pd, {}).const - 2.0) < 1e-15
    src.k.fill_(7.0)                                              # in place: version counter moves
    assert abs(src._match(src.forward, pd, {}).const - 7.0) < 1e-15
    src.k = 3.0
    assert abs(src._match(src.forward, pd, {}).const - 3.0) < 1e-15

    el = tm.ElementAssembler.from_mesh(m)
    c = {"v": 1.0}
    form = lambda gradu, gradv: c["v"] * (gradu @ gradv)
    assert el._match(form, False, {}, {}, {}).form == "laplace"
    c["v"] = 2.0                                                  # now a scaled Laplacian: not built in
    assert el._match(form, False, {}, {}, {}) is None
    fa = tm.FacetAssembler.from_mesh(m)
    sc = [1.0]
    ffn = lambda v: sc[0] * v
    assert float(fa._match(ffn, False, {}, 3).const[0]) == pytest.approx(1.0, abs=1e-15)
    sc[0] = 5.0
    assert float(fa._match(ffn, False, {}, 3).const[0]) == pytest.approx(5.0, abs=1e-15)
    # the cache is bounded: per-call lambdas do not pile up
    for i in range(100):
        node._match(lambda v, i=i: float(i) * v, pd, {})
    assert len(node._match_cache._d) <= node._match_cache.size


def test_user_form_torch_path_matches_oracle():
    # a form that is not built in runs through the generic vmap path
    m = unit_square_tri(5, 0.2, 2)

    class Aniso(tm.ElementAssembler):
        def forward(self, gradu, gradv):
            return 2.0 * gradu[0] * gradv[0] + gradu[1] * gradv[1]

    K = Aniso.from_mesh(m)()
    v, r, c, _ = fem.assemble_bilinear(m.points, m.elements(), "triangle", 2, "laplace")
    # oracle of the same anisotropic form through element_integrals' ingredients
    phi, grad, jxw, _ = fem.shape_grad_jxw(m.points, m.elements(), "triangle", 2)
    integ = 2.0 * grad[..., :, None, 0] * grad[..., None, :, 0] + grad[..., :, None, 1] * grad[..., None, :, 1]
    loc = torch.einsum("eqij,eq->eij", integ, jxw)
    ref = torch.zeros(r.numel(), dtype=torch.float64).index_add(0, torch.from_numpy(fem.pattern(m.elements().numpy(), m.n_points)[2]), loc.reshape(-1))
    assert torch.equal(K.row, r) and torch.equal(K.col, c)
    assert torch.allclose(K.values, ref, rtol=1e-13, atol=1e-15)


def test_user_load_torch_path_matches_oracle():
    m = unit_square_tri(6, 0.2, 3)
    Src = tm.func_node_assembler(lambda x: torch.sin(x[..., 0]) + x[..., 1])
    b = Src.from_mesh(m)(m.points)
    f = torch.sin(m.points[:, 0]) + m.points[:, 1]
    # f is linear in y but not in x: compare with the oracle's interpolated form only on the linear part
    Lin = tm.func_node_assembler(lambda x: x[..., 1])
    b2 = Lin.from_mesh(m)(m.points)
    ref = fem.assemble_linear(m.points, m.elements(), "triangle", 2, m.points[:, 1])
    assert torch.allclose(b2, ref, rtol=1e-13, atol=1e-16)
    assert b.shape == (m.n_points,)


def test_sparse_matrix_api():
    edata = torch.stack([torch.ones(2, 2), 2 * torch.ones(2, 2)]).double()
    M = tm.SparseMatrix.from_block_coo(edata, torch.tensor([0, 1]), torch.tensor([0, 1]), (2, 2))
    assert torch.equal(M.to_dense(), torch.tensor([[1., 1, 0, 0], [1, 1, 0, 0], [0, 0, 2, 2], [0, 0, 2, 2]]).double())
    x = torch.arange(4.0).double()
    assert torch.allclose(M @ x, M.to_dense() @ x)
    assert M.has_same_layout(M.layout_hash) and (M + M).has_same_layout(M)
    assert torch.equal(M.T.to_dense(), M.to_dense().T)


def test_condenser_cpu_known_answers():
    # tests/operator/test_condenser.py:48-138 (reference)
    row = torch.tensor([0, 0, 1, 1, 1, 2, 2, 2, 3, 3])
    col = torch.tensor([0, 1, 0, 1, 2, 1, 2, 3, 2, 3])
    val = torch.tensor([2., -1., -1., 2., -1., -1., 2., -1., -1., 2.]).double()
    K = tm.SparseMatrix(val, row, col, (4, 4))
    f = torch.tensor([1., 0., 0., 1.]).double()
    mask = torch.tensor([True, False, False, True])
    cd = tm.Condenser(mask, torch.tensor([1.0, 2.0]).double())
    Ki, fi = cd(K, f)
    assert Ki.shape == (2, 2) and torch.equal(fi, torch.tensor([1., 2.]).double())
    u = cd.recover(torch.tensor([5., 6.]).double())
    assert torch.equal(u, torch.tensor([1., 5., 6., 2.]).double())
    cd2 = tm.Condenser(mask, torch.tensor([99.0, 99.0]))
    assert torch.equal(cd2.restrict(f), torch.tensor([0., 0.]).double())
    assert torch.equal(cd2.prolong(torch.tensor([1., 2.]).double()), torch.tensor([0., 1., 2., 0.]).double())
    assert cd2.is_inner_dof is None


def test_cg_solves_condensed_poisson_cpu():
    d = np.load(ROOT / "tests" / "golden" / "cfg1_tri20.npz")
    Ki = tm.SparseMatrix(torch.from_numpy(d["Kii"]), torch.from_numpy(d["Kii_row"]),
                         torch.from_numpy(d["Kii_col"]), (361, 361))
    fi = torch.from_numpy(d["f_i"])
    u = Ki.solve(fi, rtol=1e-12)
    ref = torch.linalg.solve(Ki.to_dense(), fi)
    assert torch.allclose(u, ref, rtol=1e-9, atol=1e-12)


def test_row_centric_summation_reproduces_oracle_on_cpu():
    """Emulates the fused P1 kernel's bookkeeping (incidence lists + packed
    column positions + per-entry walk) in numpy and checks it against the
    oracle's index_add_ assembly."""
    m = unit_cube_tet(3, 0.15, 4)
    conn = m.elements()
    pat = Pattern.build(conn, m.n_points)
    cp = pat.colpos(conn).numpy().view(np.uint32)
    loc = fem.element_integrals(m.points, conn, "tetra", 1, "laplace").numpy()   # [E,4,4]
    rp, ip, inc = pat.row_ptr.numpy(), pat.inc_ptr.numpy(), pat.inc.numpy()
    vals = np.zeros(pat.nnz)
    for r in range(m.n_points):
        for k in range(rp[r], rp[r + 1]):
            p = k - rp[r]
            acc = 0.0
            for t in range(ip[r], ip[r + 1]):
                e, i = divmod(int(inc[t]), 4)
                for j in range(4):
                    if (int(cp[t]) >> (8 * j)) & 0xFF == p:
                        acc += loc[e, i, j]
            vals[k] = acc
    ref, r, c, _ = fem.assemble_bilinear(m.points, conn, "tetra", 1, "laplace")
    assert np.array_equal(vals, ref.numpy())   # same summation order -> bit-identical
    # the per-entry incidence masks select exactly the contributing incidences
    mask = pat._entry_mask.numpy().view(np.uint32)
    vals2 = np.zeros(pat.nnz)
    for r_ in range(m.n_points):
        for k in range(rp[r_], rp[r_ + 1]):
            p, acc, mm = k - rp[r_], 0.0, int(mask[k])
            while mm:
                t = ip[r_] + (mm & -mm).bit_length() - 1
                mm &= mm - 1
                e, i = divmod(int(inc[t]), 4)
                j = [jj for jj in range(4) if (int(cp[t]) >> (8 * jj)) & 0xFF == p][0]
                acc += loc[e, i, j]
            vals2[k] = acc
    assert np.array_equal(vals2, ref.numpy())
    # the plan covers every row exactly once, in order
    bp = pat.plan(Pattern.ROWS_CAP, node=True, max_rows=Pattern.ROWS_MAX).numpy()
    assert bp[0] == 0 and bp[-1] == m.n_points and np.all(np.diff(bp) > 0)
    assert np.diff(bp).max() <= Pattern.ROWS_MAX


# ------------------------------------------------------------------ facets ---
def _facet_mesh(kind, order=1, mask="all"):
    from tensormesh_amd.mesh import lognormal_coefficient, quadratic_arrays, unit_cube_tet_arrays, unit_square_tri_arrays
    if kind == "tri":
        p, c, bm = unit_square_tri_arrays(5, 0.2, 1)
        et = "triangle" if order == 1 else "triangle6"
    else:
        p, c, bm = unit_cube_tet_arrays(2, 0.1, 1)
        et = "tetra" if order == 1 else "tetra10"
    if order == 2:
        p, c, bm = quadratic_arrays(p, c, bm)
    if mask == "face":
        bm = p[:, 0] <= 1e-12
    m = tm.Mesh(torch.from_numpy(p), {et: torch.from_numpy(c)}, torch.from_numpy(bm))
    return m, et


@pytest.mark.parametrize("kind,order", [("tri", 1), ("tri", 2), ("tet", 1), ("tet", 2)])
def test_facet_topology_and_torch_path_match_oracle(kind, order):
    from oracle import fem
    m, et = _facet_mesh(kind, order)
    c = m.elements()
    elem, fac = fem.facet_select(c.numpy(), m.boundary_mask.numpy(), et)

    class Load(tm.FacetAssembler):
        def forward(self, v, f):
            return f * v

    A = Load.from_mesh(m, quadrature_order=2)
    assert A._topo.elem.tolist() == elem.tolist() and A._topo.local.tolist() == fac.tolist()
    f = torch.sin(3 * m.points[:, 0]) + m.points[:, 1]
    b = A(point_data={"f": f})
    ref = fem.assemble_facet_linear(m.points, c, m.boundary_mask, et, 2, f_nodal=f)
    assert torch.allclose(b, ref, rtol=1e-12, atol=1e-14)

    class Traction(tm.FacetAssembler):
        def forward(self, v):
            t = torch.tensor([0.25, -1.0, 0.5][:m.points.shape[1]], dtype=v.dtype)
            return v[:, None] * t[None, :]

    bt = Traction.from_assembler(A)()
    tvec = torch.tensor([0.25, -1.0, 0.5][:m.points.shape[1]], dtype=torch.float64)
    reft = fem.assemble_facet_linear(m.points, c, m.boundary_mask, et, 2, const=tvec)
    assert torch.allclose(bt, reft, rtol=1e-12, atol=1e-14)


def test_facet_torch_path_custom_integrand_uses_gradients():
    """forward(v, gradv, x) = (|x|^2 + gradv[:, 0]) v-shaped: checked against
    the oracle's facet geometry (values, gradients, FxW) directly."""
    from oracle import fem
    m, et = _facet_mesh("tri", 2)

    class Custom(tm.FacetAssembler):
        def forward(self, v, gradv, x):
            return (x * x).sum() * v + gradv[:, 0]

    b = Custom.from_mesh(m, quadrature_order=3)()
    elem, fac = fem.facet_select(m.elements().numpy(), m.boundary_mask.numpy(), et)
    phi, grad, fxw = fem.facet_geometry(m.points, m.elements(), elem, fac, et, 3)
    dofs = m.elements()[torch.from_numpy(elem)]
    xq = torch.einsum("fbd,fqb->fqd", m.points[dofs], phi)
    integrand = (xq * xq).sum(-1)[..., None] * phi + grad[..., 0]
    ref = torch.zeros(m.n_points, dtype=torch.float64).index_add(
        0, dofs.reshape(-1), torch.einsum("fqb,fq->fb", integrand, fxw).reshape(-1))
    assert torch.allclose(b, ref, rtol=1e-12, atol=1e-14)


def test_facet_builtin_recognition():
    from tensormesh_amd.facet import match_facet
    pd = {"f": torch.zeros(5), "g": torch.zeros(5, 3)}
    assert match_facet(lambda v, f: f * v, pd, 4).key == "f"
    mg = match_facet(lambda v, g: v[:, None] * g[None, :], pd, 4)
    assert mg.key == "g" and mg.n_comp == 3
    mc = match_facet(lambda v: 2.5 * v, pd, 4)
    assert mc.key is None and abs(float(mc.const[0]) - 2.5) < 1e-14
    mt = match_facet(lambda v: v[:, None] * torch.tensor([0.0, -1.0], dtype=v.dtype), pd, 4)
    assert mt.n_comp == 2 and torch.allclose(mt.const, torch.tensor([0.0, -1.0], dtype=torch.float64))
    assert match_facet(lambda v, f: f * f * v, pd, 4) is None
    assert match_facet(lambda v: v * v, pd, 4) is None


# --------------------------------------------------------- J2 plasticity ---
_J2 = dict(E=200.0, nu=0.3, sig0=0.25, H=1.0)


@pytest.mark.parametrize("etype", ["triangle", "tetra10"])
def test_j2_host_torch_path_matches_oracle(etype):
    """On host tensors J2Plasticity.energy runs the generic vmap path with the
    reference's element_energy and update_state its torch restatement; both
    against the oracle, with history given as {etype: tensor}."""
    order = 2 if etype == "tetra10" else 1
    m = unit_square_tri(2, 0.1, 3, order=order) if etype == "triangle" else unit_cube_tet(1, 0.1, 3, order=order)
    model = tm.J2Plasticity.from_mesh(m, quadrature_order=2, **_J2)
    h = model.history[etype]
    E = m.elements().shape[0]
    Q = fem.quadrature(etype, 2)[0].shape[0]
    assert h["eps_p"].shape == (E, Q, 3, 3) and h["alpha"].shape == (E, Q)
    g = torch.Generator().manual_seed(1)
    ep = 1e-4 * torch.randn(E, Q, 3, 3, generator=g, dtype=torch.float64)
    ep = 0.5 * (ep + ep.transpose(-1, -2))
    al = 1e-4 * torch.rand(E, Q, generator=g, dtype=torch.float64)
    u = (6e-4 * torch.randn(m.points.shape, generator=g, dtype=torch.float64)).requires_grad_(True)
    e = model.energy(point_data={"displacement": u}, element_data={"eps_p_n": {etype: ep}, "alpha_n": {etype: al}})
    ref = fem.j2_energy(m.points, m.elements(), etype, 2, u.detach(), ep, al, **_J2)
    assert abs(e.item() - ref.item()) <= 1e-12 * abs(ref.item())
    e.backward()
    res_e, _ = fem.j2_element_derivatives(m.points, m.elements(), etype, 2, u.detach(), ep, al, **_J2)
    D = m.points.shape[1]
    res = torch.zeros(m.n_points, D, dtype=torch.float64).index_add(0, m.elements().reshape(-1),
                                                                     res_e.reshape(-1, D))
    assert torch.allclose(u.grad, res, rtol=1e-10, atol=1e-13)
    h["eps_p"], h["alpha"] = ep.clone(), al.clone()
    model.update_state(u.detach())
    ep1, al1 = fem.j2_update(m.points, m.elements(), etype, 2, u.detach(), ep, al, **_J2)
    assert torch.allclose(model.history[etype]["eps_p"], ep1, rtol=1e-12, atol=1e-18)
    assert torch.allclose(model.history[etype]["alpha"], al1, rtol=1e-12, atol=1e-18)


def test_j2_material_object_and_native_requires_device():
    class Steel:
        E, nu, sigma_y, H = 200e9, 0.3, 250e6, 1e9

    m = unit_cube_tet(1, 0.0, 0)
    model = tm.J2Plasticity.from_mesh(m, material=Steel())
    assert (model.sig0, model.H) == (250e6, 1e9)
    assert abs(model.bulk - (model.lam + 2 / 3 * model.mu)) < 1e-3
    with pytest.raises(nat.NativeError):
        model.residual(torch.zeros_like(m.points))


# ------------------------------------------------------- contact energies ---
@pytest.mark.parametrize("kind,order", [("tri", 2), ("tet", 1)])
def test_penalty_contact_torch_path_matches_oracle(kind, order):
    """PenaltyContact.energy on host tensors (the generic vmapped path with
    the reference's example density) against the oracle, for a deformed
    configuration pressing into a tilted plane; FxW of the construction
    configuration."""
    m, et = _facet_mesh(kind, order)
    D = m.points.shape[1]
    nrm = [0.2, 1.0] if D == 2 else [0.1, 1.0, -0.2]
    model = tm.PenaltyContact.from_mesh(m, quadrature_order=2, kappa=3.0, normal=nrm, offset=0.05)
    g = torch.Generator().manual_seed(4)
    x = m.points + 0.02 * torch.randn(m.points.shape, generator=g, dtype=torch.float64)
    x[:, 1] -= 0.1
    e = model.energy(points=x)
    s = math.sqrt(sum(v * v for v in nrm))
    ref = fem.contact_energy(m.points, x, m.elements(), m.boundary_mask, et, 2, 3.0, [v / s for v in nrm], 0.05)
    assert ref.item() > 0
    assert abs(e.item() - ref.item()) <= 1e-12 * ref.item()


def test_contact_assembler_generic_arguments():
    """A user density with per-element data, a scalar and interpolated nodal
    data (the reference's broadcast rules), against explicit einsums."""
    m, et = _facet_mesh("tri", 1)
    elem, fac = fem.facet_select(m.elements().numpy(), m.boundary_mask.numpy(), et)
    phi, _, fxw = fem.facet_geometry(m.points, m.elements(), elem, fac, et, 2)

    class Robin(tm.ContactAssembler):
        def element_energy(self, x, u, k, c):
            return k * c * (u * u).sum() + x[0]

    model = Robin.from_mesh(m, quadrature_order=2)
    u = torch.sin(3 * m.points)
    k = torch.linspace(1.0, 2.0, m.elements().shape[0], dtype=torch.float64)
    e = model.energy(point_data={"u": u}, element_data={"k": {et: k}}, scalar_data={"c": 0.5})
    dofs = m.elements()[torch.from_numpy(elem)]
    uq = torch.einsum("sbd,sqb->sqd", u[dofs], phi)
    xq = torch.einsum("sbd,sqb->sqd", m.points[dofs], phi)
    dens = k[torch.from_numpy(elem)][:, None] * 0.5 * (uq * uq).sum(-1) + xq[..., 0]
    assert torch.allclose(e, (dens * fxw).sum(), rtol=1e-13)
    with pytest.raises(ValueError):
        model.energy(func=lambda y: y.sum())


# ------------------------------------------- quadrilaterals and hexahedra ---
@pytest.mark.parametrize("etype", ["quad", "quad9", "hexahedron", "hexahedron27"])
def test_tensor_cells_torch_path_and_facets_match_oracle(etype):
    """Host tensors on quadrilateral / hexahedral meshes: the generic torch
    path of ElementAssembler (diffusion with element kappa) and FacetAssembler
    (f * v) against the oracle."""
    from tensormesh_amd.mesh import unit_cube_hex, unit_square_quad
    order = 2 if etype in ("quad9", "hexahedron27") else 1
    m = (unit_square_quad(3, 0.2, 1, order=order) if etype.startswith("quad")
         else unit_cube_hex(2, 0.2, 1, order=order))
    assert m.default_element_type == etype
    kappa = torch.linspace(0.5, 2.0, m.elements().shape[0], dtype=torch.float64)
    A = tm.DiffusionElementAssembler.from_mesh(m, quadrature_order=3)
    A.reference_path = True            # built-in forms otherwise run only on the device
    K = A(element_data={"kappa": kappa})
    Ke = fem.element_integrals(m.points, m.elements(), etype, 3, "diffusion_elem", kappa=kappa)
    row, col, eids = fem.pattern(m.elements().numpy(), m.n_points)
    ref = fem._index_add(Ke.reshape(-1), torch.from_numpy(eids), row.size)
    assert np.array_equal(K.row.numpy(), row) and np.array_equal(K.col.numpy(), col)
    assert torch.allclose(K.values, ref, rtol=1e-12, atol=1e-14 * ref.abs().max().item())

    class Load(tm.FacetAssembler):
        def forward(self, v, f):
            return f * v

    f = torch.sin(2 * m.points[:, 0]) + m.points[:, 1]
    b = Load.from_mesh(m, quadrature_order=2)(point_data={"f": f})
    want = fem.assemble_facet_linear(m.points, m.elements(), m.boundary_mask, etype, 2, f_nodal=f)
    assert torch.allclose(b, want, rtol=1e-12, atol=1e-14 * want.abs().max().item())


@pytest.mark.parametrize("etype", ["triangle10", "tetra20"])
def test_cubic_cells_torch_path_matches_oracle(etype):
    from tensormesh_amd.mesh import unit_cube_tet, unit_square_tri
    m = unit_square_tri(2, 0.15, 1, order=3) if etype == "triangle10" else unit_cube_tet(1, 0.1, 1, order=3)
    assert m.default_element_type == etype
    A = tm.LaplaceElementAssembler.from_mesh(m, quadrature_order=4)
    A.reference_path = True
    K = A()
    Ke = fem.element_integrals(m.points, m.elements(), etype, 4, "laplace")
    row, col, eids = fem.pattern(m.elements().numpy(), m.n_points)
    ref = fem._index_add(Ke.reshape(-1), torch.from_numpy(eids), row.size)
    assert torch.allclose(K.values, ref, rtol=1e-12, atol=1e-13 * ref.abs().max().item())


# ------------------------------------------------------------------ tuner ---
def test_tuner_cache_roundtrip_and_disabled_default(tmp_path, monkeypatch):
    from tensormesh_amd import tune
    path = tmp_path / "tune.json"
    t = tune.Tuner(path)
    t.put("dev|op|abc", {"choice": "rows", "ms": {"rows": 0.1}})
    t2 = tune.Tuner(path)
    assert t2.get("dev|op|abc")["choice"] == "rows"
    monkeypatch.setenv("TENSORMESH_AMD_TUNE", "0")
    t3 = tune.Tuner(tmp_path / "other.json")
    calls = []
    assert t3.choose("k", {"a": lambda: calls.append("a"), "b": lambda: calls.append("b")}) == "a"
    assert calls == []          # disabled: nothing is timed
    assert t3.choose("k", {"only": lambda: calls.append("x")}) == "only"


def test_pcg_chunk_rule():
    from tensormesh_amd.tune import pcg_chunk
    assert pcg_chunk(None) == 32 and pcg_chunk(0) == 32
    assert pcg_chunk(100) == 8 and pcg_chunk(2697) == 128 and pcg_chunk(700) == 32
    for n in (1, 50, 300, 1000, 5000, 10 ** 6):
        c = pcg_chunk(n)
        assert c in (8, 16, 32, 64, 128) and c % 2 == 0


def test_cg_warns_when_the_budget_runs_out():
    """Host CG on a small SPD system: converged solves are silent, a starved
    iteration budget warns (torch_solve.py:125-126)."""
    import warnings as _w
    from tensormesh_amd.solve import cg
    m = unit_square_tri(6, 0.1, 2)
    A = tm.LaplaceElementAssembler.from_mesh(m)
    A.reference_path = True
    K = A()
    Ki, fi = tm.Condenser(m.boundary_mask)(K, torch.ones(m.n_points, dtype=torch.float64))
    with _w.catch_warnings():
        _w.simplefilter("error")
        x, it, rel = cg(Ki, fi, rtol=1e-10, check_every=1)
    assert rel <= 1e-10
    with pytest.warns(UserWarning, match="did not converge"):
        cg(Ki, fi, rtol=1e-12, max_iter=2, check_every=1)


def test_cg_refinement_reaches_the_residual_rounding_level():
    """refine=1 solves A d = b - A x once more: on a loosely stopped solve the
    error against the dense solution drops by orders of magnitude, the
    reported residual is the true one and the iterations are the total."""
    from tensormesh_amd.solve import cg
    m = unit_square_tri(12, 0.2, 3)
    A = tm.DiffusionElementAssembler.from_mesh(m)
    A.reference_path = True
    E = m.elements().shape[0]
    K = A(element_data={"kappa": torch.from_numpy(lognormal_coefficient(E, 3))})
    Ki, fi = tm.Condenser(m.boundary_mask)(K, torch.ones(m.n_points, dtype=torch.float64))
    ref = torch.linalg.solve(Ki.to_dense(), fi)
    x0, it0, _ = cg(Ki, fi, rtol=1e-6, check_every=1)
    x1, it1, rel1 = cg(Ki, fi, rtol=1e-6, check_every=1, refine=1)
    e0 = ((x0 - ref).norm() / ref.norm()).item()
    e1 = ((x1 - ref).norm() / ref.norm()).item()
    assert e1 < 1e-3 * e0
    assert it1 > it0
    true_rel = ((fi - Ki.to_dense() @ x1).norm() / fi.norm()).item()
    assert abs(rel1 - true_rel) <= 1e-3 * true_rel + 1e-18
    # differentiable: the adjoint solve is refined as well
    v = Ki.values.clone().requires_grad_(True)
    Kg = tm.SparseMatrix(v, Ki.row_indices, Ki.col_indices, Ki.shape)
    u = Kg.solve(fi, rtol=1e-6, refine=2)
    u.sum().backward()
    lam = torch.linalg.solve(Ki.to_dense(), torch.ones_like(fi))
    g_ref = -lam[Ki.row_indices] * ref[Ki.col_indices]
    assert torch.allclose(v.grad, g_ref, rtol=1e-8, atol=1e-12 * g_ref.abs().max().item())


@pytest.mark.parametrize("etype,H", [("tetra10", 3), ("triangle6", 2), ("tetra10", 1)])
def test_symmetric_pair_lists_reproduce_the_full_reduction(etype, H):
    """Pattern.sym_pairs on the CPU: summing each upper entry once from the
    symmetric slot list (bit 31 = read the block transposed) and mirroring it
    gives, bit for bit, the plain reduction of full symmetric element blocks
    (the paired form of tm_segmented_reduce_pairs, emulated in numpy)."""
    m = unit_cube_tet(2, 0.1, 1, order=2) if etype == "tetra10" else unit_square_tri(3, 0.1, 1, order=2)
    conn = m.elements()
    E, B = conn.shape
    pat = Pattern.build(conn, m.n_points)
    seg = pat.seg_slot.long()
    e, r = seg // (B * B), seg % (B * B)
    i, j = r // B, r % B
    a, b = torch.minimum(i, j), torch.maximum(i, j)
    P = B * (B + 1) // 2
    packed = e * P + a * B - a * (a - 1) // 2 + (b - a)
    sym = torch.where(i <= j, packed, packed - 2 ** 31).to(torch.int32)
    pat._sym_slot = sym                      # tm_symmetric_segments, restated
    entry, entry_t, sp, ss = (t.numpy().astype(np.int64) for t in pat.sym_pairs())
    # random element blocks with K_e[j][i] = K_e[i][j]^T exactly
    rng = np.random.default_rng(0)
    K = rng.standard_normal((E, B, B, H, H))
    K = 0.5 * (K + K.transpose(0, 2, 1, 4, 3))
    slot = K.reshape(E * B * B, H * H)
    iu, ju = np.triu_indices(B)              # row-major pairs i <= j: the packed order
    packed_slot = K[:, iu, ju].reshape(E * P, H * H)
    # plain reduction: entry k sums its slots left to right
    sp0, ss0 = pat.seg_ptr.numpy(), pat.seg_slot.numpy()
    full = np.zeros((pat.nnz, H * H))
    for k in range(pat.nnz):
        for t in range(sp0[k], sp0[k + 1]):
            full[k] += slot[ss0[t]]
    # paired reduction
    tr = np.arange(H * H).reshape(H, H).T.reshape(-1)
    out = np.full((pat.nnz, H * H), np.nan)
    for k in range(entry.size):
        acc = np.zeros(H * H)
        for t in range(sp[k], sp[k + 1]):
            raw = int(ss[t])
            s, flip = raw & 0x7FFFFFFF, raw < 0
            acc += packed_slot[s][tr] if flip else packed_slot[s]
        out[entry[k]] = acc
        if entry_t[k] != entry[k]:
            out[entry_t[k]] = acc[tr]
    assert not np.isnan(out).any(), "every entry is written"
    assert np.array_equal(out, full)
    # contribution order (slot_dest, tm_slot_dest_build restated): the element
    # kernels write packed block s at slot dest[s] (transposed when flagged);
    # every segment then sums consecutive slots -- the same values in the same order
    assert ss.size == E * P
    dest = np.empty(E * P, np.int64)
    dest[ss & 0x7FFFFFFF] = np.arange(ss.size)
    flip = ss < 0
    contig = np.empty((E * P, H * H))
    contig[dest] = packed_slot
    contig[dest[ss[flip] & 0x7FFFFFFF]] = packed_slot[ss[flip] & 0x7FFFFFFF][:, tr]
    out2 = np.full((pat.nnz, H * H), np.nan)
    for k in range(entry.size):
        acc = np.zeros(H * H)
        for t in range(sp[k], sp[k + 1]):
            acc += contig[t]
        out2[entry[k]] = acc
        if entry_t[k] != entry[k]:
            out2[entry_t[k]] = acc[tr]
    assert np.array_equal(out2, full)
    rows = np.repeat(np.arange(m.n_points), np.diff(pat.row_ptr.numpy()))
    cols = pat.col_idx.numpy()
    assert (rows[entry] <= cols[entry]).all()
    assert np.array_equal(rows[entry_t], cols[entry]) and np.array_equal(cols[entry_t], rows[entry])


def test_workspace_bytes():
    """tm_workspace_bytes (host arithmetic): slot buffers of the element and
    node stages and the J2 energy workspace, with its argument errors."""
    L = nat.lib()
    mv = nat.MeshView(nat.CELL["tetra10"], nat.TM_F64, 7, 30, None, None)
    assert L.tm_workspace_bytes(0, ctypes.byref(mv), 9) == 7 * 10 * 10 * 9 * 8
    assert L.tm_workspace_bytes(1, ctypes.byref(mv), 3) == 7 * 10 * 3 * 8
    assert L.tm_workspace_bytes(2, ctypes.byref(mv), 0) == (7 + nat.J2_ENERGY_PARTS) * 8
    mv32 = nat.MeshView(nat.CELL["hexahedron27"], nat.TM_F32, 5, 30, None, None)
    assert L.tm_workspace_bytes(0, ctypes.byref(mv32), 1) == 5 * 27 * 27 * 4
    assert L.tm_workspace_bytes(3, ctypes.byref(mv), 9) == 7 * 55 * 9 * 8   # packed upper blocks
    assert L.tm_workspace_bytes(4, ctypes.byref(mv), 1) == -1     # TM_ERR_ARG
    assert L.tm_workspace_bytes(0, ctypes.byref(mv), 0) == -1


def test_symmetric_slots_default_policy():
    """symmetric_slots=None stores symmetric blocks for the D x D block forms
    (the callers pass True there) and full blocks for scalar forms (False);
    an explicit True / False overrides either."""
    m = unit_square_tri(2, 0.0, 0, order=2)
    A = tm.LaplaceElementAssembler.from_mesh(m, quadrature_order=2)
    assert A.symmetric_slots is None
    assert A._symmetric_blocks(True) is True and A._symmetric_blocks(False) is False
    A.symmetric_slots = False
    assert A._symmetric_blocks(True) is False
    A.symmetric_slots = True
    assert A._symmetric_blocks(False) is True


def test_host_pattern_builder_under_address_and_ub_sanitizers(tmp_path):
    """The host pattern builder (tm_pattern.cpp, tm_common.cpp) built with
    AddressSanitizer and UndefinedBehaviorSanitizer (g++, host code only) and
    driven over small meshes, a repeated vertex, an empty mesh, out-of-range
    node ids, negative sizes and an over-long row (tests/native)."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not available")
    exe = tmp_path / "pattern_asan"
    src = ROOT / "tensormesh_amd" / "csrc"
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", f"-I{ROOT / 'include'}", f"-I{src}",
           str(ROOT / "tests" / "native" / "pattern_driver.cpp"), str(src / "tm_pattern.cpp"),
           str(src / "tm_common.cpp"), "-o", str(exe), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    # verify_asan_link_order=0: the process environment may preload other
    # libraries ahead of the sanitizer runtime
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.stdout + r.stderr)[-3000:]
