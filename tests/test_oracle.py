"""Pins of the CPU oracle (oracle/fem.py) against facts that do not come
from the oracle itself: closed forms, invariants, brute force and the
known answers of the reference's own tests."""
import itertools
import math

import numpy as np
import pytest
import torch

from oracle import fem
from tensormesh_amd.mesh import quadratic_arrays, unit_cube_tet_arrays, unit_square_tri_arrays

torch.set_default_dtype(torch.float64)


def _dense(v, r, c, shape):
    return torch.sparse_coo_tensor(torch.stack([r, c]), v, shape).to_dense()


# ---------------------------------------------------------------- pattern
@pytest.mark.parametrize("n,kind", [(3, "tri"), (2, "tet"), (1, "tet10")])
def test_pattern_bruteforce(n, kind):
    if kind == "tri":
        p, c, _ = unit_square_tri_arrays(n, 0.1, 3)
    elif kind == "tet":
        p, c, _ = unit_cube_tet_arrays(n, 0.1, 3)
    else:
        p, c, _ = quadratic_arrays(*unit_cube_tet_arrays(n, 0.0, 3))
    row, col, eids = fem.pattern(c, p.shape[0])
    pairs = sorted({(int(a), int(b)) for e in c for a in e for b in e})
    assert list(zip(row.tolist(), col.tolist())) == pairs
    B = c.shape[1]
    for e in range(c.shape[0]):
        for i in range(B):
            for j in range(B):
                k = eids[e * B * B + i * B + j]
                assert (row[k], col[k]) == (c[e, i], c[e, j])


def test_pattern_sizes_of_the_workloads():
    p, c, bnd = unit_square_tri_arrays(20, 0.2, 0)
    assert c.shape == (800, 3) and p.shape == (441, 2) and bnd.sum() == 80
    assert fem.pattern(c, 441)[0].size == 2921
    p, c, _ = unit_cube_tet_arrays(64, 0.15, 0)
    assert c.shape == (1572864, 4) and p.shape[0] == 274625
    assert fem.pattern(c, p.shape[0])[0].size == 4018753


# ---------------------------------------------------------- layout hash
# Digest prefixes of the reference's own construction (scipy int32 edges,
# SHA-256 of row bytes + col bytes) reproduced independently of this
# repository's code: the 20x20 triangle workload and the 4x4 split square.
_REFERENCE_DIGEST_PREFIX = {20: "f28f12ec", 4: "45799f"}


def _packed_digest(pairs, fmt):
    import hashlib
    import struct
    rows = [a for a, _ in pairs]
    cols = [b for _, b in pairs]
    return hashlib.sha256(struct.pack(f"<{len(rows)}{fmt}", *rows)
                          + struct.pack(f"<{len(cols)}{fmt}", *cols)).hexdigest()


@pytest.mark.parametrize("n", [4, 20])
def test_layout_hash_known_answers_and_encoding(n):
    p, c, _ = unit_square_tri_arrays(n, 0.2 if n == 20 else 0.0, 0)
    edges = fem.reference_edges(torch.from_numpy(c), p.shape[0])
    # scipy keeps 32-bit indices at this size; the digest hashes them as such
    assert edges.dtype == torch.int32
    h = fem.layout_hash(edges[0], edges[1])
    assert h.startswith(_REFERENCE_DIGEST_PREFIX[n])
    # independent encoding: brute-force sorted unique pairs packed as
    # little-endian int32 by struct (no numpy / scipy / torch byte path)
    pairs = sorted({(int(a), int(b)) for e in c for a in e for b in e})
    assert h == _packed_digest(pairs, "i")
    # the int64 digest of the same pairs differs (what a wrong dtype gives)
    assert h != _packed_digest(pairs, "q")


def test_layout_hash_block_and_condensed_are_int64():
    p, c, _ = unit_cube_tet_arrays(1, 0.0, 0)
    edges = fem.reference_edges(torch.from_numpy(c), p.shape[0])
    r, cc = fem.block_coo_indices(edges[0], edges[1], 3)
    assert r.dtype == torch.int64 and cc.dtype == torch.int64
    pairs = sorted({(int(a), int(b)) for e in c for a in e for b in e})
    blk = [(3 * a + i, 3 * b + j) for a, b in pairs for i in range(3) for j in range(3)]
    assert fem.layout_hash(r, cc) == _packed_digest(blk, "q")
    # condensed blocks carry the Condenser's int64 local numbering
    mask = torch.zeros(p.shape[0], dtype=torch.bool)
    mask[0] = True
    out = fem.condense(torch.ones(edges.shape[1]), edges[0].long(), edges[1].long(), p.shape[0], mask)
    _, kr, kc, _ = out["K_ii"]
    assert kr.dtype == torch.int64
    loc = {g: k for k, g in enumerate(i for i in range(p.shape[0]) if i != 0)}
    inner = [(loc[a], loc[b]) for a, b in pairs if a != 0 and b != 0]
    assert fem.layout_hash(kr, kc) == _packed_digest(inner, "q")


# ------------------------------------------------------ reference element
# degree of exactness of each rule of the reference's tables; its tetrahedron
# "order 5" (14 points), "order 6" (15 points) and "order 7" (24 points)
# rules reach degrees 4, 5 and 6
_EXACT = {("triangle", o): o for o in range(1, 8)}
_EXACT.update({("tetra", o): d for o, d in ((1, 1), (2, 2), (3, 3), (4, 4), (5, 4), (6, 5), (7, 6))})


@pytest.mark.parametrize("etype,order,vol", [("triangle", o, 0.5) for o in range(1, 8)] +
                         [("tetra", o, 1 / 6) for o in range(1, 8)])
def test_quadrature_weights_and_exactness(etype, order, vol):
    w, q = fem.quadrature(etype, order)
    assert abs(w.sum().item() - vol) < 1e-12
    dim = q.shape[1]
    deg = _EXACT[(etype, order)]
    # exact integral of x^a y^b (z^c) over the reference simplex: a! b! c! / (d + a + b + c)!
    worst_above = 0.0
    for e in itertools.product(range(deg + 2), repeat=dim):
        if sum(e) > deg + 1:
            continue
        exact = math.prod(math.factorial(k) for k in e) / math.factorial(dim + sum(e))
        approx = (w * torch.prod(q ** torch.tensor(e, dtype=torch.float64), 1)).sum().item()
        if sum(e) <= deg:
            assert abs(approx - exact) < 1e-12, (e, approx, exact)
        else:
            worst_above = max(worst_above, abs(approx - exact))
    assert worst_above > 1e-9          # the degree is not higher than stated


@pytest.mark.parametrize("etype", ["triangle", "triangle6", "tetra", "tetra10", "triangle10", "tetra20"])
def test_basis_kronecker_and_partition_of_unity(etype):
    nodes = fem.lagrange_nodes(etype)
    phi = fem.shape_values(etype, nodes)
    assert np.allclose(phi, np.eye(nodes.shape[0]), atol=1e-12)
    pts = np.random.default_rng(0).random((7, nodes.shape[1])) * 0.3
    assert np.allclose(fem.shape_values(etype, pts).sum(1), 1.0)
    assert np.allclose(fem.shape_ref_grads(etype, pts).sum(2), 0.0, atol=1e-12)


def test_p2_triangle_values_match_reference_test():
    # closed form of tests/element/test_transformation.py:66-98 (reference)
    w, q = fem.quadrature("triangle", 4)
    q = q.numpy()
    xi, eta = q[:, 0], q[:, 1]
    exp = np.zeros((q.shape[0], 6))
    exp[:, 0] = (1 - xi - eta) * (1 - 2 * xi - 2 * eta)
    exp[:, 1] = xi * (2 * xi - 1)
    exp[:, 2] = eta * (2 * eta - 1)
    exp[:, 5] = 4 * xi * (1 - xi - eta)
    exp[:, 3] = 4 * xi * eta
    exp[:, 4] = 4 * eta * (1 - xi - eta)
    assert np.allclose(fem.shape_values("triangle6", q), exp, atol=1e-13)


def test_product_tables_equal_oracle_vandermonde():
    from tensormesh_amd import element as el
    for etype in ("triangle", "triangle6", "tetra", "tetra10"):
        w, q = el.quadrature(etype, 2)
        assert np.allclose(el.shape_values(etype, q), fem.shape_values(etype, q), atol=1e-13)
        assert np.allclose(el.shape_ref_grads(etype, q), fem.shape_ref_grads(etype, q), atol=1e-12)


# -------------------------------------------------------------- geometry
def test_jacobian_determinant_is_measure():
    p, c, _ = unit_square_tri_arrays(5, 0.2, 1)
    P, C = torch.from_numpy(p), torch.from_numpy(c)
    _, grad, jxw, _ = fem.shape_grad_jxw(P, C, "triangle", 1)
    x = p[c]
    area = 0.5 * np.abs((x[:, 1, 0] - x[:, 0, 0]) * (x[:, 2, 1] - x[:, 0, 1])
                        - (x[:, 2, 0] - x[:, 0, 0]) * (x[:, 1, 1] - x[:, 0, 1]))
    assert np.allclose(jxw[:, 0].numpy(), area, rtol=1e-13)
    p, c, _ = unit_cube_tet_arrays(3, 0.15, 1)
    P, C = torch.from_numpy(p), torch.from_numpy(c)
    _, grad, jxw, _ = fem.shape_grad_jxw(P, C, "tetra", 2)
    x = p[c]
    vol = np.abs(np.einsum("ei,ei->e", x[:, 1] - x[:, 0], np.cross(x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]))) / 6
    assert np.allclose(jxw.sum(1).numpy(), vol, rtol=1e-13)
    # gradients of a linear function are exact
    a = np.array([0.3, -1.2, 2.0])
    u = torch.from_numpy(p @ a + 0.7)
    gu = torch.einsum("eb,eqbd->eqd", u[C], grad)
    assert torch.allclose(gu, torch.tensor(a).expand_as(gu), atol=1e-11)


# ---------------------------------------------------------- bilinear forms
def test_laplace_patch_test_and_properties():
    p, c, bnd = unit_square_tri_arrays(8, 0.2, 5)
    P, C = torch.from_numpy(p), torch.from_numpy(c)
    K = _dense(*fem.assemble_bilinear(P, C, "triangle", 2, "laplace"))
    assert torch.allclose(K, K.T, atol=1e-14)
    assert K.sum(1).abs().max() < 1e-12
    assert torch.linalg.eigvalsh(K).min() > -1e-10
    u = P @ torch.tensor([0.4, -0.9]) + 0.1
    assert (K @ u)[~torch.from_numpy(bnd)].abs().max() < 1e-12


def test_diffusion_element_closed_form():
    p, c, _ = unit_cube_tet_arrays(2, 0.15, 2)
    P, C = torch.from_numpy(p), torch.from_numpy(c)
    kap = torch.rand(c.shape[0]) + 0.5
    loc = fem.element_integrals(P, C, "tetra", 1, "diffusion_elem", kappa=kap)
    x = p[c]
    for e in range(c.shape[0]):
        X = np.stack([x[e, 1] - x[e, 0], x[e, 2] - x[e, 0], x[e, 3] - x[e, 0]])
        G = np.linalg.inv(X)           # columns = grad(lambda_1..3)
        g = np.concatenate([-G.sum(1, keepdims=True), G], 1)
        vol = abs(np.linalg.det(X)) / 6
        assert np.allclose(loc[e].numpy(), kap[e].item() * vol * g.T @ g, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("etype,D", [("triangle", 2), ("tetra", 3)])
def test_mass_closed_form(etype, D):
    gen = unit_square_tri_arrays if D == 2 else unit_cube_tet_arrays
    p, c, _ = gen(3, 0.1, 4)
    P, C = torch.from_numpy(p), torch.from_numpy(c)
    loc = fem.element_integrals(P, C, etype, 2, "mass")
    _, _, jxw, _ = fem.shape_grad_jxw(P, C, etype, 1)
    vol = jxw[:, 0]
    exact = vol[:, None, None] / ((D + 1) * (D + 2)) * (torch.ones(D + 1, D + 1) + torch.eye(D + 1))
    assert torch.allclose(loc, exact, rtol=1e-12)
    M = _dense(*fem.assemble_bilinear(P, C, etype, 2, "mass"))
    assert abs(M.sum().item() - 1.0) < 1e-12


def test_voigt_known_answers():
    # tests/functional/test_elasticity.py:19-84 (reference)
    B = fem.voigt_B(torch.tensor([1., 2., 3.]))
    assert torch.equal(B, torch.tensor([[1., 0, 0], [0, 2, 0], [0, 0, 3], [0, 3, 2], [3, 0, 1], [2, 1, 0]]))
    B2 = fem.voigt_B2(torch.tensor([1., 2.]))
    assert torch.equal(B2, torch.tensor([[1., 0], [0, 2], [2, 1]]))
    C = fem.voigt_C(70.0, 0.3, 3)
    mu, lam = 70 / 2.6, 70 * 0.3 / (1.3 * 0.4)
    exp = torch.zeros(6, 6)
    exp[:3, :3] = lam
    exp[[0, 1, 2], [0, 1, 2]] = lam + 2 * mu
    exp[[3, 4, 5], [3, 4, 5]] = mu
    assert torch.allclose(C, exp)


def test_elasticity_rigid_body_modes():
    p, c, _ = quadratic_arrays(*unit_cube_tet_arrays(2, 0.1, 6))
    P, C = torch.from_numpy(p), torch.from_numpy(c)
    v, r, cc, shape = fem.assemble_bilinear(P, C, "tetra10", 2, "linear_elasticity", E=2.0, nu=0.25)
    K = _dense(v, r, cc, shape)
    assert torch.allclose(K, K.T, atol=1e-12)
    n = p.shape[0]
    modes = []
    for d in range(3):
        t = torch.zeros(n, 3)
        t[:, d] = 1
        modes.append(t.reshape(-1))
    for a, b in ((0, 1), (1, 2), (0, 2)):
        t = torch.zeros(n, 3)
        t[:, a], t[:, b] = -P[:, b], P[:, a]
        modes.append(t.reshape(-1))
    for m in modes:
        assert (K @ m).abs().max() < 1e-11 * K.abs().max()


# ----------------------------------------------------- projector / blocks
def test_reduce_project_bruteforce():
    g = torch.Generator().manual_seed(0)
    for n_src, n_dst, h in ((100, 10, 1), (1000, 100, 8)):
        dst = torch.randint(0, n_dst, (n_src,), generator=g)
        x = torch.randn(n_src, h, generator=g)
        out = fem.reduce_project(x, dst, n_dst)
        ref = torch.zeros(n_dst, h)
        for s in range(n_src):
            ref[dst[s]] += x[s]
        assert torch.allclose(out, ref, atol=1e-12)


def test_block_coo_known_answer():
    # tests/sparse/test_block_coo.py:10-25 (reference)
    edata = torch.stack([torch.ones(2, 2), 2 * torch.ones(2, 2)])
    v, r, c = fem.block_coo(edata, torch.tensor([0, 1]), torch.tensor([0, 1]))
    dense = _dense(v, r, c, (4, 4))
    assert torch.equal(dense, torch.tensor([[1., 1, 0, 0], [1, 1, 0, 0], [0, 0, 2, 2], [0, 0, 2, 2]]))


# ------------------------------------------------------------------ loads
def test_load_closed_form_and_integral():
    p, c, _ = unit_square_tri_arrays(6, 0.2, 9)
    P, C = torch.from_numpy(p), torch.from_numpy(c)
    f = torch.sin(P[:, 0]) + P[:, 1] ** 2
    b = fem.assemble_linear(P, C, "triangle", 2, f)
    # exact for P1 f_h * phi (quadratic) with the order-2 rule: |T|/12 (2 f_i + f_j + f_k)
    _, _, jxw, _ = fem.shape_grad_jxw(P, C, "triangle", 1)
    exp = torch.zeros(p.shape[0])
    for e in range(c.shape[0]):
        fe = f[C[e]]
        for i in range(3):
            exp[C[e, i]] += jxw[e, 0] / 12 * (fe.sum() + fe[i])
    assert torch.allclose(b, exp, rtol=1e-12)
    ones = fem.assemble_linear(P, C, "triangle", 2, torch.ones(p.shape[0]))
    assert abs(ones.sum().item() - 1.0) < 1e-13


# ------------------------------------------------------------- condenser
def test_condenser_known_answers():
    # tests/operator/test_condenser.py:48-100 (reference): 4x4 tridiagonal
    row = torch.tensor([0, 0, 1, 1, 1, 2, 2, 2, 3, 3])
    col = torch.tensor([0, 1, 0, 1, 2, 1, 2, 3, 2, 3])
    val = torch.tensor([2., -1., -1., 2., -1., -1., 2., -1., -1., 2.])
    f = torch.tensor([1., 0., 0., 1.])
    mask = torch.tensor([True, False, False, True])
    out = fem.condense(val, row, col, 4, mask, torch.tensor([1.0, 2.0]), f)
    kv, kr, kc, shape = out["K_ii"]
    assert shape == (2, 2)
    assert torch.equal(_dense(kv, kr, kc, shape), torch.tensor([[2., -1.], [-1., 2.]]))
    # f_i = f_inner - K_io u_o = [0 - (-1*1), 0 - (-1*2)]
    assert torch.equal(out["f_i"], torch.tensor([1., 2.]))
    u = fem.recover(torch.tensor([5., 6.]), mask, out["dvalue"])
    assert torch.equal(u, torch.tensor([1., 5., 6., 2.]))


# ------------------------------------------------------------ neo-hookean
def test_neohookean_zero_state_and_linearisation():
    p, c, _ = unit_cube_tet_arrays(1, 0.1, 2)
    P, C = torch.from_numpy(p), torch.from_numpy(c)
    u0 = torch.zeros_like(P)
    assert abs(fem.neohookean_energy(P, C, "tetra", 2, u0).item()) < 1e-14
    res, tan = fem.neohookean_element_tangent(P, C, "tetra", 2, u0, E=1.0, nu=0.3)
    assert res.abs().max() < 1e-14
    lin = fem.element_integrals(P, C, "tetra", 2, "linear_elasticity", E=1.0, nu=0.3)
    assert torch.allclose(tan, lin, atol=1e-13)


def test_neohookean_finite_differences():
    p, c, _ = unit_cube_tet_arrays(1, 0.1, 2)
    P, C = torch.from_numpy(p), torch.from_numpy(c)
    g = torch.Generator().manual_seed(3)
    u = 0.05 * torch.randn(P.shape, generator=g)
    res, tan = fem.neohookean_element_tangent(P, C, "tetra", 2, u)
    h = 1e-6
    # residual == d energy / du (element 0, dof (1, 2))
    e = 0
    def elem_energy(uu):
        return fem.neohookean_energy(P, C[e:e + 1], "tetra", 2, uu)
    du = torch.zeros_like(u)
    du[C[e, 1], 2] = h
    fd = (elem_energy(u + du) - elem_energy(u - du)) / (2 * h)
    assert abs(fd.item() - res[e, 1, 2].item()) < 1e-8
    # tangent == d residual / du
    res_p, _ = fem.neohookean_element_tangent(P, C, "tetra", 2, u + du)
    res_m, _ = fem.neohookean_element_tangent(P, C, "tetra", 2, u - du)
    fd_t = (res_p - res_m)[e] / (2 * h)               # d R[e,i,a] / d u[node1, 2]
    assert torch.allclose(fd_t, tan[e, :, 1, :, 2], atol=1e-7)


# ------------------------------------------------------------------ facets ---
def test_facet_line_rule_exact_on_monomials():
    """n-point Gauss-Legendre on [0, 1] integrates x^k exactly for k <= 2n - 1."""
    for n in (1, 2, 3, 4):
        w, p = fem.facet_quadrature("triangle", n)
        assert abs(float(w.sum()) - 1.0) < 1e-15
        for k in range(2 * n):
            assert abs(float((w * p[:, 0] ** k).sum()) - 1.0 / (k + 1)) < 1e-14


def test_facet_select_brute_force():
    """Selection against plain loops: every facet (local node set) of every
    element with a flagged node, kept iff all its nodes are flagged."""
    from tensormesh_amd.mesh import unit_cube_tet_arrays, quadratic_arrays
    for order in (1, 2):
        p, c, bm = unit_cube_tet_arrays(2, 0.1, 4)
        et = "tetra"
        if order == 2:
            p, c, bm = quadratic_arrays(p, c, bm)
            et = "tetra10"
        faces = {1: [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]}[1]
        mids = {(2, 3): 4, (1, 3): 5, (1, 2): 6, (0, 3): 7, (0, 2): 8, (0, 1): 9}
        expect = []
        for e in range(c.shape[0]):
            if not any(bm[c[e]]):
                continue
            for f, fv in enumerate(faces):
                nodes = list(fv)
                if order == 2:
                    nodes += [mids[(a, b)] for (a, b) in mids if a in fv and b in fv]
                if all(bm[c[e, k]] for k in nodes):
                    expect.append((e, f))
        elem, fac = fem.facet_select(c, bm, et)
        assert list(zip(elem.tolist(), fac.tolist())) == expect


def test_facet_square_closed_forms():
    """P1 triangles of the unit square, f = 1 and f = x over all boundary
    facets: 4 + 4 sqrt(2) h and 2 + 2 sqrt(2) h -- the reference also selects
    the two corner-square diagonals, whose end nodes are both on the boundary
    (once from each of their two triangles); one edge only: |edge| = 1,
    traction sums to t."""
    from tensormesh_amd.mesh import unit_square_tri_arrays
    n = 8
    h = 1.0 / n
    p, c, bm = unit_square_tri_arrays(n, 0.2, 3)
    P, C, BM = torch.from_numpy(p), torch.from_numpy(c), torch.from_numpy(bm)
    for q in (1, 2, 3):
        b1 = fem.assemble_facet_linear(P, C, BM, "triangle", q, const=torch.tensor(1.0))
        bx = fem.assemble_facet_linear(P, C, BM, "triangle", q, f_nodal=P[:, 0].clone())
        assert abs(float(b1.sum()) - (4 + 4 * math.sqrt(2) * h)) < 1e-13
        assert abs(float(bx.sum()) - (2 + 2 * math.sqrt(2) * h)) < 1e-13
    left = torch.from_numpy(p[:, 0] <= 1e-12)
    b = fem.assemble_facet_linear(P, C, left, "triangle", 2, const=torch.tensor(1.0))
    assert abs(float(b.sum()) - 1.0) < 1e-14
    # interior nodes of the edge get h, its two ends h / 2
    ys = p[left.numpy(), 1]
    vals = b[left].numpy()
    assert np.allclose(vals[(ys > 1e-12) & (ys < 1 - 1e-12)], h, rtol=0, atol=1e-14)
    assert np.allclose(vals[(ys <= 1e-12) | (ys >= 1 - 1e-12)], h / 2, rtol=0, atol=1e-14)
    bt = fem.assemble_facet_linear(P, C, left, "triangle", 2, const=torch.tensor([0.5, -1.0])).reshape(-1, 2)
    assert torch.allclose(bt.sum(0), torch.tensor([0.5, -1.0], dtype=torch.float64), atol=1e-14)


def test_facet_p2_simpson_and_vertex_zero():
    """P2: on an edge, f = 1 gives Simpson weights (vertex h/6, midpoint 2h/3);
    on a triangular face the quadratic vertex functions integrate to 0 and
    each midpoint function to |face|/3."""
    from tensormesh_amd.mesh import unit_square_tri_arrays, unit_cube_tet_arrays, quadratic_arrays
    n = 4
    h = 1.0 / n
    p, c, bm = quadratic_arrays(*unit_square_tri_arrays(n, 0.0, 0))
    P, C = torch.from_numpy(p), torch.from_numpy(c)
    left = torch.from_numpy(p[:, 0] <= 1e-12)
    b = fem.assemble_facet_linear(P, C, left, "triangle6", 2, const=torch.tensor(1.0)).numpy()
    ys = p[:, 1]
    onl = left.numpy()
    is_mid = np.arange(p.shape[0]) >= (n + 1) ** 2
    assert np.allclose(b[onl & is_mid], 2 * h / 3, atol=1e-14)
    inner_v = onl & ~is_mid & (ys > 1e-12) & (ys < 1 - 1e-12)
    assert np.allclose(b[inner_v], h / 3, atol=1e-14)
    p3, c3, bm3 = quadratic_arrays(*unit_cube_tet_arrays(2, 0.0, 0))
    P3, C3 = torch.from_numpy(p3), torch.from_numpy(c3)
    face = torch.from_numpy(p3[:, 0] <= 1e-12)
    b3 = fem.assemble_facet_linear(P3, C3, face, "tetra10", 2, const=torch.tensor(1.0)).numpy()
    nv = 27
    assert np.abs(b3[:nv]).max() < 1e-15
    assert abs(b3.sum() - 1.0) < 1e-14


def test_facet_cube_areas_brute_force():
    """All boundary facets of a jittered Kuhn cube (with the interior faces
    whose three vertices lie on the boundary): sum of b for f = 1 equals the
    sum of the selected triangles' areas computed with cross products; the
    face x = 0 alone has area 1 and integral of y = 1/2."""
    from tensormesh_amd.mesh import unit_cube_tet_arrays
    p, c, bm = unit_cube_tet_arrays(3, 0.1, 2)
    P, C, BM = torch.from_numpy(p), torch.from_numpy(c), torch.from_numpy(bm)
    faces = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]
    area = 0.0
    for e in range(c.shape[0]):
        for fv in faces:
            nodes = [c[e, k] for k in fv]
            if all(bm[nodes]):
                a, b_, d = p[nodes[0]], p[nodes[1]], p[nodes[2]]
                area += 0.5 * np.linalg.norm(np.cross(b_ - a, d - a))
    for q in (1, 2, 3):
        s = float(fem.assemble_facet_linear(P, C, BM, "tetra", q, const=torch.tensor(1.0)).sum())
        assert abs(s - area) < 1e-13 * max(1.0, area)
    face = torch.from_numpy(p[:, 0] <= 1e-12)
    assert abs(float(fem.assemble_facet_linear(P, C, face, "tetra", 2, const=torch.tensor(1.0)).sum()) - 1.0) < 1e-14
    y = fem.assemble_facet_linear(P, C, face, "tetra", 2, f_nodal=P[:, 1].clone())
    assert abs(float(y.sum()) - 0.5) < 1e-14


# --------------------------------------------------------- J2 plasticity ---
_J2 = dict(E=200.0, nu=0.3, sig0=0.25, H=1.0)


def _j2_state(E, Q, seed, scale=0.0):
    g = torch.Generator().manual_seed(seed)
    ep = scale * torch.randn(E, Q, 3, 3, generator=g)
    ep = 0.5 * (ep + ep.transpose(-1, -2))
    ep = ep - ep.diagonal(dim1=-2, dim2=-1).sum(-1)[..., None, None] / 3.0 * torch.eye(3)
    al = scale * torch.rand(E, Q, generator=g)
    return ep, al


@pytest.mark.parametrize("etype", ["triangle", "tetra"])
def test_j2_elastic_limit_is_linear_elasticity(etype):
    """Below yield (sig0 -> inf) the incremental potential is the linear
    elastic energy 1/2 u^T K u, and its Hessian the linear-elasticity
    element matrices (plane strain in 2D: the strain is padded with zeros)."""
    p, c, _ = (unit_square_tri_arrays(2, 0.1, 5) if etype == "triangle" else unit_cube_tet_arrays(1, 0.1, 5))
    P, C = torch.from_numpy(p), torch.from_numpy(c)
    Q = fem.quadrature(etype, 2)[0].shape[0]
    ep, al = _j2_state(C.shape[0], Q, 1)
    u = 0.01 * torch.randn(P.shape, generator=torch.Generator().manual_seed(2))
    prm = dict(_J2, sig0=1e9)
    lin = fem.element_integrals(P, C, etype, 2, "linear_elasticity", E=prm["E"], nu=prm["nu"])  # [E,B,B,D,D]
    ue = u[C]
    e_lin = 0.5 * torch.einsum("eia,eijab,ejb->", ue, lin, ue)
    e_j2 = fem.j2_energy(P, C, etype, 2, u, ep, al, **prm)
    assert abs(e_j2.item() - e_lin.item()) < 1e-12 * abs(e_lin.item())
    res, tan = fem.j2_element_derivatives(P, C, etype, 2, u, ep, al, **prm)
    assert torch.allclose(tan, lin, rtol=1e-12, atol=1e-10)
    assert torch.allclose(res, torch.einsum("eijab,ejb->eia", lin, ue), rtol=1e-11, atol=1e-12)


def test_j2_simple_shear_closed_form():
    """Homogeneous simple shear u = (gamma y, 0, 0) on the unit cube: the
    trial deviator has norm gamma / sqrt 2, so the energy is
    V (mu gamma^2 / 2 - (2 mu + 2H/3) dg^2 / 2) with
    dg = max(sqrt(2) mu gamma - sqrt(2/3) sig0, 0) / (2 mu + 2H/3)."""
    p, c, _ = unit_cube_tet_arrays(2, 0.15, 7)
    P, C = torch.from_numpy(p), torch.from_numpy(c)
    Q = fem.quadrature("tetra", 2)[0].shape[0]
    ep, al = torch.zeros(C.shape[0], Q, 3, 3), torch.zeros(C.shape[0], Q)
    mu = _J2["E"] / (2 * (1 + _J2["nu"]))
    den = 2 * mu + 2.0 / 3.0 * _J2["H"]
    for gamma in (1e-4, 2e-3, 1e-2):
        u = torch.zeros_like(P)
        u[:, 0] = gamma * P[:, 1]
        dg = max(math.sqrt(2.0) * mu * gamma - math.sqrt(2.0 / 3.0) * _J2["sig0"], 0.0) / den
        want = mu * gamma ** 2 / 2 - 0.5 * den * dg ** 2
        got = fem.j2_energy(P, C, "tetra", 2, u, ep, al, **_J2).item()
        assert abs(got - want) < 1e-12 * abs(want), (gamma, got, want)
    assert dg > 0  # the last case yields


@pytest.mark.parametrize("etype", ["triangle6", "tetra"])
def test_j2_finite_differences_plastic(etype):
    """Residual = central difference of the energy and tangent = central
    difference of the residual, with a strain large enough that most
    quadrature points are past yield and a nonzero prior plastic state."""
    if etype == "tetra":
        p, c, _ = unit_cube_tet_arrays(1, 0.1, 2)
    else:
        p, c, _ = quadratic_arrays(*unit_square_tri_arrays(1, 0.1, 2))
    P, C = torch.from_numpy(p), torch.from_numpy(c)
    Q = fem.quadrature(etype, 2)[0].shape[0]
    ep, al = _j2_state(C.shape[0], Q, 4, 1e-3)
    u = 0.01 * torch.randn(P.shape, generator=torch.Generator().manual_seed(5))
    G, _, _ = fem.j2_grad_u(P, C, etype, 2, u)
    mu, _ = fem.j2_moduli(_J2["E"], _J2["nu"])
    res, tan = fem.j2_element_derivatives(P, C, etype, 2, u, ep, al, **_J2)
    e, node = 0, 1
    h = 1e-7
    for a in range(P.shape[1]):
        du = torch.zeros_like(u)
        du[C[e, node], a] = h
        fd = (fem.j2_energy(P, C[e:e + 1], etype, 2, u + du, ep[e:e + 1], al[e:e + 1], **_J2)
              - fem.j2_energy(P, C[e:e + 1], etype, 2, u - du, ep[e:e + 1], al[e:e + 1], **_J2)) / (2 * h)
        assert abs(fd.item() - res[e, node, a].item()) < 1e-6 * res[e].abs().max().item()
        rp, _ = fem.j2_element_derivatives(P, C, etype, 2, u + du, ep, al, **_J2)
        rm, _ = fem.j2_element_derivatives(P, C, etype, 2, u - du, ep, al, **_J2)
        fd_t = (rp - rm)[e] / (2 * h)
        assert torch.allclose(fd_t, tan[e, :, node, :, a], rtol=1e-5, atol=1e-5 * tan[e].abs().max().item())
    # the case is plastic: most points have a positive trial yield value
    eps_tr = fem._j2_strain3(G) - ep
    dev = eps_tr - eps_tr.diagonal(dim1=-2, dim2=-1).sum(-1)[..., None, None] / 3 * torch.eye(3)
    f_tr = 2 * mu * torch.linalg.norm(dev, dim=(-2, -1)) - math.sqrt(2 / 3) * (_J2["sig0"] + _J2["H"] * al)
    assert (f_tr > 0).float().mean() > 0.5


def test_j2_update_returns_to_the_yield_surface():
    """After the radial return, a point that yielded lies on the updated
    yield surface (2 mu |dev(eps - eps_p')| = sqrt(2/3)(sig0 + H alpha')),
    the plastic strain stays trace-free, alpha grows by sqrt(2/3) |d eps_p|,
    and points below yield keep their state."""
    p, c, _ = unit_cube_tet_arrays(2, 0.1, 9)
    P, C = torch.from_numpy(p), torch.from_numpy(c)
    Q = fem.quadrature("tetra", 2)[0].shape[0]
    ep, al = _j2_state(C.shape[0], Q, 6, 1e-4)
    u = 0.0002 * torch.randn(P.shape, generator=torch.Generator().manual_seed(8))
    mu, _ = fem.j2_moduli(_J2["E"], _J2["nu"])
    ep1, al1 = fem.j2_update(P, C, "tetra", 2, u, ep, al, **_J2)
    G, _, _ = fem.j2_grad_u(P, C, "tetra", 2, u)
    dev = lambda x: x - x.diagonal(dim1=-2, dim2=-1).sum(-1)[..., None, None] / 3 * torch.eye(3)
    f0 = 2 * mu * torch.linalg.norm(dev(fem._j2_strain3(G) - ep), dim=(-2, -1)) \
        - math.sqrt(2 / 3) * (_J2["sig0"] + _J2["H"] * al)
    yielded = f0 > 0
    assert 0 < int(yielded.sum()) < yielded.numel()
    f1 = 2 * mu * torch.linalg.norm(dev(fem._j2_strain3(G) - ep1), dim=(-2, -1)) \
        - math.sqrt(2 / 3) * (_J2["sig0"] + _J2["H"] * al1)
    assert f1[yielded].abs().max() < 1e-12
    assert torch.equal(ep1[~yielded], ep[~yielded]) and torch.equal(al1[~yielded], al[~yielded])
    assert ep1.diagonal(dim1=-2, dim2=-1).sum(-1).abs().max() < 1e-15
    d = torch.linalg.norm(ep1 - ep, dim=(-2, -1))
    assert torch.allclose(al1 - al, math.sqrt(2 / 3) * d, rtol=1e-12, atol=1e-18)
    # the energy after the update is the elastic energy of the new state
    e1 = fem.j2_energy(P, C, "tetra", 2, u, ep1, al1, **_J2)
    e_el = fem.j2_energy(P, C, "tetra", 2, u, ep1, al1, **dict(_J2, sig0=1e9))
    assert abs(e1.item() - e_el.item()) < 1e-9 * abs(e_el.item())


# ------------------------------------------------------- penalty contact ---
@pytest.mark.parametrize("kind", ["tri", "tet"])
def test_contact_energy_brute_force(kind):
    """Rigid translation by -h along y against the plane y = 0: facets between
    the rows y = 0 and y = h penetrate by h - y (a polynomial on each
    facet), the rest not at all.  Against per-facet adaptive quadrature of
    kappa/2 (h - y)_+^2 over the straight facet (scipy quad / dblquad with the
    facet's own measure), independent of the shape-function tables."""
    import scipy.integrate as si
    n = 3 if kind == "tri" else 2
    if kind == "tri":
        p, c, bm = unit_square_tri_arrays(n, 0.0, 0)
        et = "triangle"
    else:
        p, c, bm = unit_cube_tet_arrays(n, 0.0, 0)
        et = "tetra"
    P, C, BM = torch.from_numpy(p), torch.from_numpy(c), torch.from_numpy(bm)
    h, kappa = 1.0 / n, 7.0
    x = P.clone()
    x[:, 1] -= h
    nrm = [0.0, 1.0] if kind == "tri" else [0.0, 1.0, 0.0]
    got = fem.contact_energy(P, x, C, BM, et, 2, kappa, nrm, 0.0).item()
    elem, fac = fem.facet_select(c, bm, et)
    loc = fem.facet_local_nodes(et)
    want = 0.0
    for e, f in zip(elem, fac):
        v = p[c[e, loc[f]]]
        pen = lambda y: 0.5 * kappa * max(h - y, 0.0) ** 2
        if kind == "tri":
            L = np.linalg.norm(v[1] - v[0])
            want += L * si.quad(lambda t: pen((1 - t) * v[0, 1] + t * v[1, 1]), 0, 1, epsabs=1e-14)[0]
        else:
            A = np.linalg.norm(np.cross(v[1] - v[0], v[2] - v[0]))
            g = lambda t, s: pen(v[0, 1] + s * (v[1, 1] - v[0, 1]) + t * (v[2, 1] - v[0, 1]))
            want += A * si.dblquad(g, 0, 1, 0, lambda s: 1 - s, epsabs=1e-14)[0]
    assert want > 0
    assert abs(got - want) < 1e-12 * want, (got, want)
    # lifted clear of the plane: no energy
    assert fem.contact_energy(P, P + 1.0, C, BM, et, 2, kappa, nrm, 0.0).item() == 0.0


def test_contact_finite_differences():
    """Gradient = central difference of the energy, Hessian = central
    difference of the gradient, for a tilted plane cutting through facets."""
    p, c, bm = quadratic_arrays(*unit_square_tri_arrays(2, 0.1, 3))
    P, C, BM = torch.from_numpy(p), torch.from_numpy(c), torch.from_numpy(bm)
    nrm = [0.3 / math.hypot(0.3, 1.0), 1.0 / math.hypot(0.3, 1.0)]
    g = torch.Generator().manual_seed(2)
    x = P + 0.01 * torch.randn(P.shape, generator=g)
    x[:, 1] -= 0.2
    res, hes, elem, fac = fem.contact_facet_derivatives(P, x, C, BM, "triangle6", 3, 5.0, nrm, 0.05)
    assert res.abs().max() > 0
    # a facet in contact, and one of its nodes with a nonzero gradient
    fi = int(torch.argmax(res.abs().amax((1, 2))))
    b = int(torch.argmax(res[fi].abs().amax(1)))
    node = C[elem[fi], b]
    hstep = 1e-6
    E = lambda xx: fem.contact_energy(P, xx, C, BM, "triangle6", 3, 5.0, nrm, 0.05)
    grad_full = torch.zeros_like(P).index_add(0, C[torch.from_numpy(elem)].reshape(-1), res.reshape(-1, 2))
    for a in range(2):
        dx = torch.zeros_like(x)
        dx[node, a] = hstep
        fd = (E(x + dx) - E(x - dx)) / (2 * hstep)
        assert abs(fd.item() - grad_full[node, a].item()) < 1e-7 * grad_full.abs().max().item()
        rp = fem.contact_facet_derivatives(P, x + dx, C, BM, "triangle6", 3, 5.0, nrm, 0.05)[0]
        rm = fem.contact_facet_derivatives(P, x - dx, C, BM, "triangle6", 3, 5.0, nrm, 0.05)[0]
        fd_h = (rp - rm)[fi] / (2 * hstep)
        assert torch.allclose(fd_h, hes[fi, :, b, :, a], atol=1e-6 * hes[fi].abs().max().item())


# ------------------------------------------- quadrilaterals and hexahedra ---
_TENSOR = ["quad", "quad9", "hexahedron", "hexahedron27"]


@pytest.mark.parametrize("etype", ["quad", "hexahedron"])
def test_tensor_rule_exact_on_monomials(etype):
    """The n-point-per-axis rule integrates x^a y^b (z^c) exactly on [0, 1]^D
    for every exponent up to 2n - 1."""
    dim = 2 if etype == "quad" else 3
    for n in (1, 2, 3, 4):
        w, p = fem.quadrature(etype, n)
        assert w.numel() == n ** dim and abs(w.sum().item() - 1.0) < 1e-14
        for e in itertools.product(range(2 * n), repeat=dim):
            got = (w * torch.prod(p ** torch.tensor(e, dtype=torch.float64), 1)).sum().item()
            want = 1.0 / np.prod([k + 1 for k in e])
            assert abs(got - want) < 1e-13, (n, e)


@pytest.mark.parametrize("etype", _TENSOR)
def test_tensor_basis_kronecker_and_partition(etype):
    nodes = fem.lagrange_nodes(etype)
    assert np.allclose(fem.shape_values(etype, nodes), np.eye(nodes.shape[0]), atol=1e-12)
    pts = np.random.default_rng(0).random((7, nodes.shape[1]))
    assert np.allclose(fem.shape_values(etype, pts).sum(1), 1.0, atol=1e-12)
    assert np.allclose(fem.shape_ref_grads(etype, pts).sum(2), 0.0, atol=1e-11)
    # the product code's closed-form tensor Lagrange basis agrees
    from tensormesh_amd import element as el
    assert np.allclose(el.shape_values(etype, pts), fem.shape_values(etype, pts), atol=1e-12)
    assert np.allclose(el.shape_ref_grads(etype, pts), fem.shape_ref_grads(etype, pts), atol=1e-11)
    assert np.array_equal(el.tensor_nodes(etype), nodes)


@pytest.mark.parametrize("etype", ["quad", "hexahedron"])
def test_unit_cell_q1_matrices_closed_form(etype):
    """Q1 stiffness and mass of the unit cell as Kronecker products of the
    1-D element matrices k = [[1,-1],[-1,1]], m = [[1/3,1/6],[1/6,1/3]]
    (x fastest in the lexicographic node order)."""
    dim = 2 if etype == "quad" else 3
    P = torch.tensor(fem.lagrange_nodes(etype))
    C = torch.arange(2 ** dim)[None]
    k = np.array([[1.0, -1.0], [-1.0, 1.0]])
    m = np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]])
    if dim == 2:
        K = np.kron(m, k) + np.kron(k, m)
        M = np.kron(m, m)
    else:
        K = np.kron(np.kron(m, m), k) + np.kron(np.kron(m, k), m) + np.kron(np.kron(k, m), m)
        M = np.kron(np.kron(m, m), m)
    got_k = fem.element_integrals(P, C, etype, 2, "laplace")[0].numpy()
    got_m = fem.element_integrals(P, C, etype, 2, "mass")[0].numpy()
    assert np.allclose(got_k, K, atol=1e-14) and np.allclose(got_m, M, atol=1e-15)


@pytest.mark.parametrize("etype,qo", [("quad", 2), ("quad9", 3), ("hexahedron", 2), ("hexahedron27", 3)])
def test_tensor_patch_test_and_rigid_modes(etype, qo):
    """On jittered grids: the assembled Laplace matrix annihilates a linear
    field in every interior row (the rule is exact for adj(J)^T grad phi),
    and the element elasticity matrices annihilate rigid-body motions."""
    from tensormesh_amd.mesh import unit_cube_hex_arrays, unit_square_quad_arrays
    order = 2 if etype in ("quad9", "hexahedron27") else 1
    if etype.startswith("quad"):
        p, c, bm = unit_square_quad_arrays(3, 0.2, 5, order=order)
    else:
        p, c, bm = unit_cube_hex_arrays(2, 0.2, 5, order=order)
    P, C = torch.from_numpy(p), torch.from_numpy(c)
    Ke = fem.element_integrals(P, C, etype, qo, "laplace")
    row, col, eids = fem.pattern(c, p.shape[0])
    vals = fem._index_add(Ke.reshape(-1), torch.from_numpy(eids), row.size)
    K = _dense(vals, torch.from_numpy(row), torch.from_numpy(col), (p.shape[0], p.shape[0]))
    lin = P @ torch.arange(1.0, P.shape[1] + 1, dtype=torch.float64) + 0.5
    r = K @ lin
    assert r[~torch.from_numpy(bm)].abs().max() < 1e-12
    Ee = fem.element_integrals(P, C, etype, qo, "linear_elasticity", E=1.0, nu=0.3)   # [E,B,B,D,D]
    D, B = P.shape[1], C.shape[1]
    KE = Ee.permute(0, 1, 3, 2, 4).reshape(C.shape[0], B * D, B * D)
    X = P[C]                                                                  # [E,B,D]
    modes = [torch.zeros_like(X) + torch.eye(D, dtype=torch.float64)[a] for a in range(D)]
    if D == 2:
        modes.append(torch.stack([-X[..., 1], X[..., 0]], -1))
    else:
        for a, b in ((0, 1), (1, 2), (0, 2)):
            rot = torch.zeros_like(X)
            rot[..., a], rot[..., b] = -X[..., b], X[..., a]
            modes.append(rot)
    for mo in modes:
        assert (KE @ mo.reshape(C.shape[0], -1, 1)).abs().max() < 1e-12


@pytest.mark.parametrize("etype", _TENSOR)
def test_tensor_facet_loads_closed_form(etype):
    """On grids with straight boundaries (only interior nodes jittered) the
    selected facets are exactly the boundary: c * v sums to the perimeter /
    surface area, and f * v with f = x + 2y (+ 3z) to the boundary integral of f
    (6 on the unit square, 3 + 6 + 9 = 18 on the unit cube)."""
    from tensormesh_amd.mesh import unit_cube_hex_arrays, unit_square_quad_arrays
    order = 2 if etype in ("quad9", "hexahedron27") else 1
    if etype.startswith("quad"):
        p, c, bm = unit_square_quad_arrays(3, 0.2, 2, order=order)
    else:
        p, c, bm = unit_cube_hex_arrays(2, 0.2, 2, order=order)
    P, C, BM = torch.from_numpy(p), torch.from_numpy(c), torch.from_numpy(bm)
    dim = P.shape[1]
    one = fem.assemble_facet_linear(P, C, BM, etype, 2, const=torch.tensor(1.0))
    assert abs(one.sum().item() - (4.0 if dim == 2 else 6.0)) < 1e-13
    f = P @ torch.arange(1.0, dim + 1, dtype=torch.float64)
    lin = fem.assemble_facet_linear(P, C, BM, etype, 2, f_nodal=f)
    assert abs(lin.sum().item() - (6.0 if dim == 2 else 18.0)) < 1e-12
    elem, fac = fem.facet_select(c, bm, etype)
    n_cells = 3 if dim == 2 else 2
    assert len(elem) == (4 * n_cells if dim == 2 else 6 * n_cells * n_cells)


# ------------------------------------------------------------ cubic simplices ---
@pytest.mark.parametrize("etype", ["triangle10", "tetra20"])
def test_p3_closed_form_basis_patch_test_and_rigid_modes(etype):
    """The product's barycentric P3 basis equals the oracle's Vandermonde
    inversion; on jittered P3 meshes the assembled Laplace matrix annihilates
    a linear field in interior rows (the order-4 rule is exact for the degree-4
    integrand of straight-sided cells) and the element elasticity matrices
    annihilate rigid motions; the P3 mass matrix of the reference cell with the
    degree-6 rule integrates 1 exactly (sum = area / volume)."""
    from tensormesh_amd import element as el
    from tensormesh_amd.mesh import cubic_arrays
    tri = etype == "triangle10"
    pts = np.random.default_rng(3).random((9, 2 if tri else 3)) * 0.3
    assert np.allclose(el.shape_values(etype, pts), fem.shape_values(etype, pts), atol=1e-12)
    assert np.allclose(el.shape_ref_grads(etype, pts), fem.shape_ref_grads(etype, pts), atol=1e-11)
    assert np.array_equal(el.facet_nodes(etype), fem.facet_local_nodes(etype))
    p, c, bm = cubic_arrays(*(unit_square_tri_arrays(3, 0.2, 4) if tri else unit_cube_tet_arrays(2, 0.15, 4)))
    P, C = torch.from_numpy(p), torch.from_numpy(c)
    Ke = fem.element_integrals(P, C, etype, 4, "laplace")
    row, col, eids = fem.pattern(c, p.shape[0])
    vals = fem._index_add(Ke.reshape(-1), torch.from_numpy(eids), row.size)
    K = _dense(vals, torch.from_numpy(row), torch.from_numpy(col), (p.shape[0], p.shape[0]))
    lin = P @ torch.arange(1.0, P.shape[1] + 1, dtype=torch.float64) - 0.25
    interior = torch.ones(p.shape[0], dtype=torch.bool)
    interior[torch.from_numpy(np.unique(c[:, fem.facet_local_nodes(etype)][
        torch.from_numpy(bm)[torch.from_numpy(c[:, fem.facet_local_nodes(etype)])].all(-1).numpy()]))] = False
    assert (K @ lin)[interior].abs().max() < 1e-11
    Ee = fem.element_integrals(P, C, etype, 4, "linear_elasticity", E=1.0, nu=0.3)
    D, B = P.shape[1], C.shape[1]
    KE = Ee.permute(0, 1, 3, 2, 4).reshape(C.shape[0], B * D, B * D)
    X = P[C]
    rot = torch.zeros_like(X)
    rot[..., 0], rot[..., 1] = -X[..., 1], X[..., 0]
    for mo in [torch.zeros_like(X) + torch.eye(D, dtype=torch.float64)[a] for a in range(D)] + [rot]:
        assert (KE @ mo.reshape(C.shape[0], -1, 1)).abs().max() < 1e-11
    ref = torch.tensor(fem.lagrange_nodes(etype)[:D + 1])
    M = fem.element_integrals(ref, torch.arange(B)[None] * 0 + torch.arange(B)[None], etype, 6, "mass") \
        if False else None
    w, q = fem.quadrature(etype, 6)
    phi = torch.tensor(fem.shape_values(etype, q.numpy()))
    Mref = torch.einsum("q,qi,qj->ij", w, phi, phi)
    assert abs(Mref.sum().item() - (0.5 if tri else 1 / 6)) < 1e-12
