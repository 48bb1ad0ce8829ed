"""float64 CPU restatement of the reference assembly chain.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Each function names the reference lines it restates.  Floating-point work
is done with torch-CPU in float64 using the same operation sequence as the
reference (einsum, batched ``torch.inverse`` / ``torch.linalg.det``,
``index_add_``), so that autograd through it reproduces the reference's
gradients; integer pattern work uses numpy/scipy exactly as
``ElementAssembler.from_mesh`` does.
"""

from __future__ import annotations

import hashlib
import itertools
from typing import Dict, Optional, Tuple

import numpy as np
import scipy.sparse
import torch

from tensormesh_amd.element import TET_QUAD, TRI_QUAD  # data tables only

FAMILY = {"triangle": ("tri", 2, 1), "triangle6": ("tri", 2, 2),
          "tetra": ("tet", 3, 1), "tetra10": ("tet", 3, 2),
          "triangle10": ("tri", 2, 3), "tetra20": ("tet", 3, 3),
          "quad": ("quad", 2, 1), "quad9": ("quad", 2, 2),
          "hexahedron": ("hex", 3, 1), "hexahedron27": ("hex", 3, 2)}
TRI_EDGE = [[1, 2], [0, 2], [0, 1]]                         # element.py:1684
TET_EDGE = [[2, 3], [1, 3], [1, 2], [0, 3], [0, 2], [0, 1]]  # element.py:1949
# tensor-product cells (element.py:1811-1814, 2100-2104)
QUAD_POINTS = [[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [1.0, 1.0]]
QUAD_EDGE = [[0, 1], [0, 2], [1, 3], [2, 3]]
HEX_POINTS = [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [1.0, 1.0, 0.0],
              [0.0, 0.0, 1.0], [1.0, 0.0, 1.0], [0.0, 1.0, 1.0], [1.0, 1.0, 1.0]]
HEX_EDGE = [[0, 1], [0, 2], [0, 4], [1, 3], [1, 5], [2, 3], [2, 6], [3, 7], [4, 5], [4, 6], [5, 7], [6, 7]]
HEX_FACE = [(0, 1, 2, 3), (0, 1, 4, 5), (0, 2, 4, 6), (1, 3, 5, 7), (2, 3, 6, 7), (4, 5, 6, 7)]


# ---------------------------------------------------------------- pattern ---
def pattern(conn: np.ndarray, n_points: int):
    """Edges and slot->edge map of ``ElementAssembler.from_mesh``.

    Restates tensormesh/assemble/element_assembler.py:804-836: slot
    ``s = e*B*B + i*B + j`` pairs ``(conn[e,i], conn[e,j])``; scipy
    ``coo -> csr -> coo`` yields the sorted unique pairs (CSR order); a second
    CSR with data ``arange(nnz)`` is the lookup table for ``elem_eids``.
    Returns (row[nnz], col[nnz], eids[S]) as int64 numpy arrays.
    """
    conn = np.asarray(conn, dtype=np.int64)
    n_elem, nb = conn.shape
    elem_u = np.repeat(conn, nb, axis=1).reshape(-1)          # conn[e, i] for j fast
    elem_v = np.tile(conn, (1, nb)).reshape(-1)               # conn[e, j]
    tmp = scipy.sparse.coo_matrix((np.ones_like(elem_u), (elem_u, elem_v)),
                                  shape=(n_points, n_points)).tocsr().tocoo()
    row, col = tmp.row.astype(np.int64), tmp.col.astype(np.int64)
    eids_csr = scipy.sparse.coo_matrix((np.arange(row.size), (row, col)),
                                       shape=(n_points, n_points)).tocsr()
    eids = np.asarray(eids_csr[elem_u, elem_v]).ravel().astype(np.int64)
    return row, col, eids


def reference_edges(conn, n_points: int) -> torch.Tensor:
    """The ``edges`` buffer ``ElementAssembler.from_mesh`` builds and later
    hands to ``SparseMatrix`` as ``edges[0]`` / ``edges[1]``, at the dtype it
    really has.

    Restates element_assembler.py:811-825 (slot pairs stacked per (i, j),
    flattened, scipy ``coo -> csr -> coo``), :863 (``torch.from_numpy`` of the
    stacked scipy ``row``/``col``: scipy chooses their int dtype) and :857-876
    (``assembler.type(dtype)`` casts floating buffers only; ``edges`` keeps
    its integer dtype).
    """
    conn = torch.as_tensor(conn).long()
    n_elem, nb = conn.shape
    elem_u, elem_v = [], []
    for i in range(nb):
        for j in range(nb):
            elem_u.append(conn[:, i])
            elem_v.append(conn[:, j])
    elem_u = torch.stack(elem_u, -1).flatten().numpy().copy()
    elem_v = torch.stack(elem_v, -1).flatten().numpy().copy()
    tmp = scipy.sparse.coo_matrix((np.ones_like(elem_u), (elem_u, elem_v)),
                                  shape=(n_points, n_points)).tocsr().tocoo()
    return torch.from_numpy(np.stack([tmp.row, tmp.col], 0))


def layout_hash(row: torch.Tensor, col: torch.Tensor) -> str:
    """``SparseMatrix.__init__``'s digest (tensormesh/sparse/matrix.py:81-86):
    SHA-256 of the bytes of the row tensor it was handed followed by those of
    the col tensor, each at its own dtype."""
    row_cpu, col_cpu = row.detach().cpu(), col.detach().cpu()
    return hashlib.sha256(row_cpu.numpy().tobytes() + col_cpu.numpy().tobytes()).hexdigest()


def block_coo_indices(row: torch.Tensor, col: torch.Tensor, block_size: int):
    """Index half of ``SparseMatrix.from_block_coo`` (matrix.py:362-378), in
    the reference's own torch ops so that torch's type promotion fixes the
    dtype of the indices handed to ``SparseMatrix`` (int32 edges times an
    int64 ``arange`` give int64)."""
    row_exp = row[:, None].repeat(1, block_size * block_size)
    col_exp = col[:, None].repeat(1, block_size * block_size)
    i, j = torch.meshgrid(torch.arange(block_size), torch.arange(block_size), indexing="ij")
    return (row_exp * block_size + i.flatten()).flatten(), (col_exp * block_size + j.flatten()).flatten()


# ----------------------------------------------------- reference element ---
def line_rule(order: int):
    """lin_quadrature (quadrature.py:13-66): n-point Gauss-Legendre, weights
    halved and points mapped to [0, 1]."""
    w, p = GAUSS_LEGENDRE[order]
    return torch.tensor(w, dtype=torch.float64) / 2, (torch.tensor(p, dtype=torch.float64) + 1) / 2


def quadrature(etype: str, order: int):
    """Quadrature table (tensormesh/element/quadrature.py:91-160); for
    quadrilaterals and hexahedra the tensor product of the ``order``-point line
    rule built as quad_quadrature / hex_quadrature do (quadrature.py:163-239):
    ``torch.meshgrid(..., indexing='xy')``, flattened."""
    fam = FAMILY[etype][0]
    if fam == "quad":
        w, p = line_rule(order)
        wx, wy = torch.meshgrid(w, w, indexing="xy")
        px, py = torch.meshgrid(p, p, indexing="xy")
        return wx.flatten() * wy.flatten(), torch.stack([px.flatten(), py.flatten()], -1)
    if fam == "hex":
        w, p = line_rule(order)
        wx, wy, wz = torch.meshgrid(w, w, w, indexing="xy")
        px, py, pz = torch.meshgrid(p, p, p, indexing="xy")
        return (wx.flatten() * wy.flatten() * wz.flatten(),
                torch.stack([px.flatten(), py.flatten(), pz.flatten()], -1))
    w, p = (TRI_QUAD if fam == "tri" else TET_QUAD)[order]
    return torch.tensor(w, dtype=torch.float64), torch.tensor(p, dtype=torch.float64)


def lagrange_nodes(etype: str) -> np.ndarray:
    """Basis nodes: vertices, then edge midpoints (basis.py:639-696,750-859);
    for quadrilaterals / hexahedra also the face centres and the cell centre
    (quad_basis / hex_basis, basis.py:697-748, 861-905)."""
    fam, dim, order = FAMILY[etype]
    if fam in ("quad", "hex"):
        verts = np.array(QUAD_POINTS if fam == "quad" else HEX_POINTS)
        if order == 1:
            return verts
        edges = QUAD_EDGE if fam == "quad" else HEX_EDGE
        faces = [(0, 1, 2, 3)] if fam == "quad" else HEX_FACE
        out = [verts] + [0.5 * (verts[a] + verts[b]) for a, b in edges] + [verts[list(f)].mean(0) for f in faces]
        if fam == "hex":
            out.append(np.full(3, 0.5))
        return np.vstack(out)
    verts = np.vstack([np.zeros(dim), np.eye(dim)])
    if order == 1:
        return verts
    edges = TRI_EDGE if fam == "tri" else TET_EDGE
    if order == 2:
        return np.vstack([verts] + [0.5 * (verts[a] + verts[b]) for a, b in edges])
    # order 3: lin_linspace along each edge (k/3, k = 1, 2 from the edge's first
    # vertex), then tri_linspace of each face without its boundary (the centroid)
    faces = [(0, 1, 2)] if fam == "tri" else [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]
    out = [verts]
    for a, b in edges:
        out += [verts[a] + t * (verts[b] - verts[a]) for t in (1 / 3, 2 / 3)]
    out += [verts[list(f)].mean(0) for f in faces]
    return np.vstack(out)


def _exponents(dim: int, order: int, tensor: bool = False):
    """Monomial exponents: total degree <= order (simplices), or each
    exponent <= order (Polynomial.tens_exp, tensor-product cells)."""
    return [e for e in itertools.product(range(order + 1), repeat=dim) if tensor or sum(e) <= order]


def vandermonde_coefficients(etype: str) -> Tuple[np.ndarray, list]:
    """coef[B, M] with phi_n = sum_m coef[n, m] * monomial_m.

    Restates ``Element.get_basis_fns`` (tensormesh/element/element.py:636-659):
    ``V = monomials(basis nodes)``, ``coef = inv(V).T``.
    """
    fam, dim, order = FAMILY[etype]
    nodes = lagrange_nodes(etype)
    exps = _exponents(dim, order, fam in ("quad", "hex"))
    V = np.array([[np.prod(node ** np.array(e)) for e in exps] for node in nodes])
    return np.linalg.inv(V).T, exps


def shape_values(etype: str, pts: np.ndarray) -> np.ndarray:
    coef, exps = vandermonde_coefficients(etype)
    mono = np.array([[np.prod(p ** np.array(e)) for e in exps] for p in pts])  # [Q, M]
    return mono @ coef.T                                                          # [Q, B]


def shape_ref_grads(etype: str, pts: np.ndarray) -> np.ndarray:
    """dphi[Q, D, B] (layout of ``Polynomials.map`` in element.py:960)."""
    coef, exps = vandermonde_coefficients(etype)
    dim = pts.shape[1]
    out = np.zeros((pts.shape[0], dim, coef.shape[0]))
    for d in range(dim):
        dm = np.zeros((pts.shape[0], len(exps)))
        for m, e in enumerate(exps):
            if e[d] == 0:
                continue
            e2 = list(e)
            e2[d] -= 1
            dm[:, m] = e[d] * np.prod(pts ** np.array(e2), axis=1)
        out[:, d, :] = dm @ coef.T
    return out


def shape_grad_jxw(points: torch.Tensor, conn: torch.Tensor, etype: str, qorder: int):
    """Physical shape gradients and JxW.

    Restates ``Element.eval_shape_grad`` (tensormesh/element/element.py:956-968)
    and ``Transformation.batch_shape_grad_jxw`` (transformation.py:1265-1272):
    ``J = einsum('ebj,qib->eqij')``, ``inv = torch.inverse(J)``,
    ``grad = einsum('qib,eqji->eqbj')``, ``jxw = w * |det J|``.
    Returns (phi[Q,B], grad[E,Q,B,D], jxw[E,Q], J[E,Q,D,D]).
    """
    w, q = quadrature(etype, qorder)
    qn = q.numpy()
    phi = torch.tensor(shape_values(etype, qn), dtype=points.dtype)
    dphi = torch.tensor(shape_ref_grads(etype, qn), dtype=points.dtype)
    coords = points[conn]                                           # [E, B, D]
    J = torch.einsum("ebj,qib->eqij", coords, dphi)
    inv = torch.inverse(J)
    grad = torch.einsum("qib,eqji->eqbj", dphi, inv)
    jxw = torch.einsum("q,eq->eq", w.to(points.dtype), torch.linalg.det(J).abs())
    return phi, grad, jxw, J


# -------------------------------------------------------------- integrands ---
def voigt_B(g: torch.Tensor) -> torch.Tensor:
    """Voigt strain-displacement rows for one basis gradient g[..., 3].

    Restates ``voigt_shape_grad`` (tensormesh/functional/elasticity.py:132-200):
    rows [diag(g); skew(g, sign=False)] = [xx, yy, zz, yz, xz, xy].
    """
    z = torch.zeros_like(g[..., 0])
    g0, g1, g2 = g[..., 0], g[..., 1], g[..., 2]
    rows = [torch.stack([g0, z, z], -1), torch.stack([z, g1, z], -1), torch.stack([z, z, g2], -1),
            torch.stack([z, g2, g1], -1), torch.stack([g2, z, g0], -1), torch.stack([g1, g0, z], -1)]
    return torch.stack(rows, -2)                                    # [..., 6, 3]


def voigt_B2(g: torch.Tensor) -> torch.Tensor:
    """2-D Voigt rows [xx, yy, xy] (elasticity.py:132-200, dim 2)."""
    z = torch.zeros_like(g[..., 0])
    g0, g1 = g[..., 0], g[..., 1]
    return torch.stack([torch.stack([g0, z], -1), torch.stack([z, g1], -1),
                        torch.stack([g1, g0], -1)], -2)             # [..., 3, 2]


def voigt_C(E: float, nu: float, dim: int) -> torch.Tensor:
    """Isotropic Voigt stiffness (elasticity.py:202-284)."""
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    n = 3 if dim == 2 else 6
    C = torch.zeros(n, n, dtype=torch.float64)
    C[:dim, :dim] = lam
    for k in range(dim):
        C[k, k] = lam + 2.0 * mu
    for k in range(dim, n):
        C[k, k] = mu
    return C


def element_integrals(points: torch.Tensor, conn: torch.Tensor, etype: str, qorder: int,
                      form: str, kappa: Optional[torch.Tensor] = None,
                      E: float = 1.0, nu: float = 0.3) -> torch.Tensor:
    """Per-element integrals [E,B,B] or [E,B,B,H,H].

    Restates the vmap'd form evaluation and ``_integrate``
    (tensormesh/assemble/element_assembler.py:359-441 and :190-209) for the
    built-in forms: Laplace ``gradu @ gradv`` (builtin.py:61-62), mass
    ``u * v`` (:99-100), element-coefficient diffusion ``kappa * (gradu @
    gradv)`` with ``element_data={'kappa': [E]}`` (thermal_topology.py:48-52),
    node-coefficient diffusion with ``point_data={'kappa': [N]}``
    (forms.rst:172-184), linear elasticity ``Ba.T @ C @ Bb`` (builtin.py:172-178).
    """
    phi, grad, jxw, _ = shape_grad_jxw(points, conn, etype, qorder)
    dim = points.shape[1]
    if form == "mass":
        integrand = torch.einsum("qi,qj->qij", phi, phi)
        return torch.einsum("qij,eq->eij", integrand, jxw)
    if form == "laplace":
        integrand = torch.einsum("eqid,eqjd->eqij", grad, grad)
    elif form == "diffusion_elem":
        integrand = kappa[:, None, None, None] * torch.einsum("eqid,eqjd->eqij", grad, grad)
    elif form == "diffusion_node":
        kq = torch.einsum("eb,qb->eq", kappa[conn], phi)
        integrand = kq[:, :, None, None] * torch.einsum("eqid,eqjd->eqij", grad, grad)
    elif form == "linear_elasticity":
        C = voigt_C(E, nu, dim).to(points.dtype)
        B = voigt_B(grad) if dim == 3 else voigt_B2(grad)          # [E,Q,B,V,D]
        integrand = torch.einsum("eqiva,vw,eqjwb->eqijab", B, C, B)
    else:
        raise ValueError(form)
    return torch.einsum("eqij...,eq->eij...", integrand, jxw)


# -------------------------------------------------------------- projector ---
def reduce_project(x: torch.Tensor, indices: torch.Tensor, n_out: int) -> torch.Tensor:
    """``ReduceProjector.__call__`` (tensormesh/assemble/projector.py:95-123):
    ``zeros(n_out).index_add_(0, indices, x.reshape(S, ...))``."""
    return _index_add(x, indices, n_out)


def _index_add(x: torch.Tensor, indices: torch.Tensor, n_out: int) -> torch.Tensor:
    """Flatten the leading dims of ``x`` that cover ``indices.numel()`` slots
    and accumulate them into ``n_out`` rows in slot order."""
    S = indices.numel()
    lead = 1
    k = 0
    while lead < S:
        lead *= x.shape[k]
        k += 1
    tail = x.shape[k:]
    flat = x.reshape(S, *tail)
    out = torch.zeros(n_out, *tail, dtype=x.dtype)
    return out.index_add(0, indices, flat)


def block_coo(values: torch.Tensor, row: torch.Tensor, col: torch.Tensor):
    """``SparseMatrix.from_block_coo`` (tensormesh/sparse/matrix.py:331-381):
    value index ``edge*H*H + a*H + b``, row ``edge_u*H + a``, col ``edge_v*H + b``."""
    H = values.shape[1]
    a = torch.arange(H).repeat_interleave(H)
    b = torch.arange(H).repeat(H)
    r = (row[:, None] * H + a[None]).reshape(-1)
    c = (col[:, None] * H + b[None]).reshape(-1)
    return values.reshape(-1), r, c


def assemble_bilinear(points: torch.Tensor, conn: torch.Tensor, etype: str, qorder: int, form: str,
                      kappa: Optional[torch.Tensor] = None, E: float = 1.0, nu: float = 0.3,
                      pat=None):
    """Full ``ElementAssembler.__call__`` for one element block.

    Returns (values, row, col, shape) of the output COO; block forms are
    expanded with ``block_coo`` (element_assembler.py:211-229).
    """
    n = points.shape[0]
    if pat is None:
        pat = pattern(conn.numpy(), n)
    row, col, eids = (torch.from_numpy(a) for a in pat)
    loc = element_integrals(points, conn, etype, qorder, form, kappa, E, nu)
    vals = _index_add(loc, eids, row.numel())
    if vals.dim() == 1:
        return vals, row, col, (n, n)
    v, r, c = block_coo(vals, row, col)
    H = vals.shape[1]
    return v, r, c, (n * H, n * H)


def assemble_linear(points: torch.Tensor, conn: torch.Tensor, etype: str, qorder: int,
                    f_nodal: torch.Tensor) -> torch.Tensor:
    """Load vector of ``NodeAssembler`` with ``forward(v, f) = f * v``.

    Restates tensormesh/assemble/node_assembler.py:357-409: nodal ``f`` is
    interpolated to the quadrature points (``einsum('eb,qb->eq')``), the
    integrand ``f_h * phi`` is integrated against ``jxw`` one quadrature point
    at a time (default ``batch_size=1``) and scattered with ``index_add_`` over
    ``conn.flatten()`` (node_assembler.py:665-670).
    """
    phi, grad, jxw, _ = shape_grad_jxw(points, conn, etype, qorder)
    fq = torch.einsum("eb,qb->eq", f_nodal[conn], phi)
    elem = None
    for q in range(phi.shape[0]):
        part = torch.einsum("eqb,eq->eb", (fq[:, q:q + 1, None] * phi[None, q:q + 1, :]), jxw[:, q:q + 1])
        elem = part if elem is None else elem + part
    return torch.zeros(points.shape[0], dtype=points.dtype).index_add(0, conn.reshape(-1), elem.reshape(-1))


# ----------------------------------------------------------------- facets ---
# Gauss-Legendre n-point rules on [-1, 1] (tensormesh/element/quadrature.py:36-54)
GAUSS_LEGENDRE = {
    1: ([2.0], [0.0]),
    2: ([1.0, 1.0], [-0.57735026918962576451, 0.57735026918962576451]),
    3: ([0.55555555555555555556, 0.88888888888888888889, 0.55555555555555555556],
        [-0.77459666924148337704, 0.0, 0.77459666924148337704]),
    4: ([0.34785484513745385737, 0.65214515486254614263, 0.65214515486254614263, 0.34785484513745385737],
        [-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526]),
}
# facet vertices per cell (element.py:1684-1685 edges of the triangle, :1950 faces of the tetrahedron)
FACET_VERTS = {"tri": [(1, 2), (0, 2), (0, 1)], "tet": [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)],
               "quad": [tuple(e) for e in QUAD_EDGE], "hex": HEX_FACE}


def facet_quadrature(etype: str, order: int):
    """Reference-facet rule in facet coordinates: line facets use the n-point
    Gauss-Legendre rule mapped to [0, 1] with halved weights
    (quadrature.py:13-66, n = order); triangle facets the triangle rule
    (quadrature.py:362-376)."""
    fam = FAMILY[etype][0]
    if fam in ("tri", "quad"):
        w, p = line_rule(order)                    # facet_quadrature_2d (quadrature.py:321-364)
        return w, p[:, None]
    if fam == "hex":
        return quadrature("quad", order)           # hex_facet_quadrature (quadrature.py:410-450)
    return quadrature("triangle", order)


def facet_local_nodes(etype: str):
    """Local nodes of each facet (vertices, then the order-2 edge midpoints on
    it, then a hexahedron face's centre; basis.py:1127-1287)."""
    fam, dim, order = FAMILY[etype]
    edges = {"tri": TRI_EDGE, "tet": TET_EDGE, "quad": QUAD_EDGE, "hex": HEX_EDGE}[fam]
    n_vert = {"tri": 3, "tet": 4, "quad": 4, "hex": 8}[fam]
    out = []
    for fi, fv in enumerate(FACET_VERTS[fam]):
        on = [m for m, (a, b) in enumerate(edges) if a in fv and b in fv]
        extra = []
        if order == 2:
            extra = [n_vert + m for m in on]
            if fam == "hex":
                extra.append(n_vert + len(HEX_EDGE) + fi)
        elif order == 3:
            extra = [n_vert + 2 * m + k for m in on for k in (0, 1)]
            if fam == "tet":
                extra.append(n_vert + 2 * len(edges) + fi)
        out.append(list(fv) + extra)
    return np.array(out)


def facet_select(conn: np.ndarray, boundary_mask: np.ndarray, etype: str):
    """``FacetAssembler.from_elements`` for cells with one facet type
    (facet_assembler.py:488-527): the boundary elements are those with any
    flagged node; a facet of one of them is selected iff all its nodes are
    flagged (also an interior facet whose nodes all lie on the boundary);
    selected facets in row-major order of the [boundary element, facet] mask.
    Returns (element index, local facet index) of each selected facet."""
    conn = np.asarray(conn)
    is_be = boundary_mask[conn].any(-1)
    be_idx = np.nonzero(is_be)[0]
    faces = conn[be_idx][:, facet_local_nodes(etype)]        # [n_be, n_facet, n_node_per_facet]
    m = boundary_mask[faces].all(-1)
    e_sel, f_sel = np.nonzero(m)
    return be_idx[e_sel], f_sel


def facet_geometry(points: torch.Tensor, conn: torch.Tensor, elem: np.ndarray, fac: np.ndarray, etype: str,
                   order: int):
    """Cell basis values phi[F, Qf, B], physical gradients grad[F, Qf, B, D]
    and FxW[F, Qf] at the facet quadrature points of the selected facets.

    Restates Transformation.facet_shape_val / facet_shape_grad / FxW
    (transformation.py:516-713, 985-1097): the reference facet is mapped
    affinely onto facet f of the reference cell (element.py:972-1067), the
    cell Jacobian J[i][j] = sum_b x_b[j] dphi_b/dxi_i is evaluated there,
    the facet tangents are F = (d gamma / d r) J and FxW = w |sqrt(det(F F^T))|."""
    fam, dim, _ = FAMILY[etype]
    w, r = facet_quadrature(etype, order)
    verts = lagrange_nodes(etype)[:{"tri": 3, "tet": 4, "quad": 4, "hex": 8}[fam]]
    tabs = []
    for fv in FACET_VERTS[fam]:
        a = verts[fv[0]]
        dg = np.stack([verts[fv[k + 1]] - a for k in range(dim - 1)])   # [D-1, D]
        xi = a[None] + r.numpy() @ dg
        tabs.append((shape_values(etype, xi), shape_ref_grads(etype, xi), dg))
    phi = torch.tensor(np.stack([t[0] for t in tabs]))[torch.from_numpy(fac)]     # [F, Qf, B]
    dphi = torch.tensor(np.stack([t[1] for t in tabs]))[torch.from_numpy(fac)]    # [F, Qf, D, B]
    dg = torch.tensor(np.stack([t[2] for t in tabs]))[torch.from_numpy(fac)]      # [F, D-1, D]
    x = points[conn[torch.from_numpy(elem)]]                                       # [F, B, D]
    J = torch.einsum("fbj,fqib->fqij", x, dphi)
    T = torch.einsum("fki,fqij->fqkj", dg, J)
    detF = torch.sqrt(torch.linalg.det(T @ T.transpose(-1, -2)))
    fxw = w[None, :] * detF.abs()
    grad = torch.einsum("fqib,fqji->fqbj", dphi, torch.inverse(J))
    return phi, grad, fxw


def assemble_facet_linear(points: torch.Tensor, conn: torch.Tensor, boundary_mask: torch.Tensor, etype: str,
                          order: int, f_nodal: Optional[torch.Tensor] = None,
                          const: Optional[torch.Tensor] = None) -> torch.Tensor:
    """``FacetAssembler.__call__`` (facet_assembler.py:123-270) with the
    integrand ``f * v`` (f nodal, [N] or [N, H], interpolated at the facet
    points) or ``c * v`` (c a constant scalar or [H] vector): the integral
    ``einsum('fqb...,fq->fb...')`` against FxW, scattered with ``index_add_``
    over the selected facets' cell nodes (facet_assembler.py:520-525).
    Returns the flattened [N] or [N*H] vector."""
    elem, fac = facet_select(conn.numpy(), boundary_mask.numpy(), etype)
    phi, _, fxw = facet_geometry(points, conn, elem, fac, etype, order)
    dofs = conn[torch.from_numpy(elem)]                                             # [F, B]
    if f_nodal is not None:
        fq = torch.einsum("fb...,fqb->fq...", f_nodal[dofs], phi)                   # [F, Qf, (H)]
    else:
        fq = const.to(torch.float64).expand(phi.shape[:2] + tuple(const.shape))
    integrand = phi[..., None] * fq[:, :, None, :] if fq.dim() == 3 else phi * fq[:, :, None]
    integral = torch.einsum("fqb...,fq->fb...", integrand, fxw)
    n = points.shape[0]
    out = torch.zeros((n,) + tuple(integral.shape[2:]), dtype=torch.float64)
    return out.index_add(0, dofs.reshape(-1), integral.reshape((-1,) + tuple(integral.shape[2:]))).reshape(-1)


# -------------------------------------------------------------- condenser ---
def condense(values, row, col, n_dof, mask, dvalue=None, rhs=None):
    """``Condenser._compute_layout`` + ``__call__`` (tensormesh/operator/condense.py:182-296).

    Returns dict with K_ii (values,row,col,shape), K_io, f_i and the local
    numbering used by ``recover``.
    """
    inner = ~mask
    is_inner_edge = inner[row] & inner[col]
    is_ou2in_edge = inner[row] & mask[col]
    n_in, n_out = int(inner.sum()), int(mask.sum())
    local = torch.full((n_dof,), -1, dtype=torch.long)
    local[inner] = torch.arange(n_in)
    local[mask] = torch.arange(n_out)
    if dvalue is None:
        dvalue = torch.zeros(n_out, dtype=values.dtype)
    elif dvalue.shape[0] == n_dof:
        dvalue = dvalue[mask]
    if rhs is None:
        rhs = torch.zeros(n_dof, dtype=values.dtype)
    kio_v = values[is_ou2in_edge]
    kio_r = local[row[is_ou2in_edge]]
    kio_c = local[col[is_ou2in_edge]]
    minus = torch.zeros(n_in, dtype=values.dtype).index_add(0, kio_r, kio_v * dvalue[kio_c])
    return {
        "K_ii": (values[is_inner_edge], local[row[is_inner_edge]], local[col[is_inner_edge]], (n_in, n_in)),
        "K_io": (kio_v, kio_r, kio_c, (n_in, n_out)),
        "f_i": rhs[inner] - minus,
        "dvalue": dvalue,
    }


def recover(u_inner, mask, dvalue):
    """``Condenser.recover`` (condense.py:339-368)."""
    out = torch.zeros(mask.shape[0], dtype=u_inner.dtype)
    out[~mask] = u_inner
    out[mask] = dvalue.to(u_inner.dtype)
    return out


# ----------------------------------------------------------- neo-hookean ---
def cofactor_det(F: torch.Tensor) -> torch.Tensor:
    """det by cofactor expansion; same value as torch.linalg.det, but with a
    finite second derivative at F = I (torch.linalg.det's double backward
    returns NaN there)."""
    if F.shape[-1] == 2:
        return F[..., 0, 0] * F[..., 1, 1] - F[..., 0, 1] * F[..., 1, 0]
    return (F[..., 0, 0] * (F[..., 1, 1] * F[..., 2, 2] - F[..., 1, 2] * F[..., 2, 1])
            - F[..., 0, 1] * (F[..., 1, 0] * F[..., 2, 2] - F[..., 1, 2] * F[..., 2, 0])
            + F[..., 0, 2] * (F[..., 1, 0] * F[..., 2, 1] - F[..., 1, 1] * F[..., 2, 0]))


def neohookean_psi(F: torch.Tensor, mu: float, lam: float, det=torch.linalg.det) -> torch.Tensor:
    """psi = mu/2 (I1 - d) - mu ln J + lam/2 (ln J)^2 (builtin.py:301-314)."""
    d = F.shape[-1]
    J = det(F)
    I1 = (F * F).sum((-2, -1))
    lj = torch.log(J)
    return 0.5 * mu * (I1 - d) - mu * lj + 0.5 * lam * lj * lj


def lame(E: float, nu: float):
    return E / (2.0 * (1.0 + nu)), E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))


def neohookean_energy(points, conn, etype, qorder, u, E=1.0, nu=0.3):
    """``NeoHookeanModel.energy`` via ``ElementAssembler.energy``
    (builtin.py:316-333, element_assembler.py:631-657):
    ``graddisplacement = einsum('eb...,eqbd->eq...d', u[conn], shape_grad)``,
    ``F = I + grad``, energy ``sum(psi * jxw)``."""
    mu, lam = lame(E, nu)
    _, grad, jxw, _ = shape_grad_jxw(points, conn, etype, qorder)
    gu = torch.einsum("eba,eqbd->eqad", u[conn], grad)
    F = torch.eye(points.shape[1], dtype=u.dtype) + gu
    return (neohookean_psi(F, mu, lam) * jxw).sum()


def neohookean_element_tangent(points, conn, etype, qorder, u, E=1.0, nu=0.3):
    """Per-element residual [E,B,H] and tangent [E,B,B,H,H] by autograd.

    The reference has no in-tree tangent: its residual is ``autograd`` of
    the energy (element_assembler.py:447-659) and a Newton solve would take
    the autograd Jacobian.  This computes exactly those derivatives element by
    element with ``torch.func`` (small meshes only).
    """
    mu, lam = lame(E, nu)
    _, grad, jxw, _ = shape_grad_jxw(points, conn, etype, qorder)
    ue = u[conn]                                                    # [E,B,H]
    dim = points.shape[1]
    eye = torch.eye(dim, dtype=u.dtype)

    def e_energy(ue_, g_, w_):
        F = eye + torch.einsum("ba,qbd->qad", ue_, g_)
        return (neohookean_psi(F, mu, lam, det=cofactor_det) * w_).sum()

    res = torch.func.vmap(torch.func.grad(e_energy))(ue, grad, jxw)
    hes = torch.func.vmap(torch.func.hessian(e_energy))(ue, grad, jxw)  # [E,B,H,B,H]
    return res, hes.permute(0, 1, 3, 2, 4).contiguous()


# --------------------------------------------------------- J2 plasticity ---
def j2_moduli(E: float, nu: float):
    """mu, bulk = lam + 2/3 mu (builtin.py:436-440)."""
    mu, lam = lame(E, nu)
    return mu, lam + 2.0 / 3.0 * mu


def _j2_strain3(G: torch.Tensor) -> torch.Tensor:
    """sym(grad u), padded to 3 x 3 in 2D (builtin.py:483-488, 530-537)."""
    eps = 0.5 * (G + G.transpose(-1, -2))
    if G.shape[-1] == 2:
        eps = torch.nn.functional.pad(eps, (0, 1, 0, 1))
    return eps


def j2_psi(G, eps_p, alpha, E, nu, sig0, H):
    """Algorithmic incremental potential at quadrature points, batched over
    leading dimensions: builtin.py:464-511 (trial strain, deviator, trial
    yield value, clamp, vol + dev energy)."""
    mu, bulk = j2_moduli(E, nu)
    eps_tr = _j2_strain3(G) - eps_p
    tr = eps_tr.diagonal(dim1=-2, dim2=-1).sum(-1)
    dev = eps_tr - (tr / 3.0)[..., None, None] * torch.eye(3, dtype=G.dtype)
    nrm = torch.linalg.norm(dev, dim=(-2, -1))
    f_tr = 2.0 * mu * nrm - np.sqrt(2.0 / 3.0) * (sig0 + H * alpha)
    denom = 2.0 * mu + (2.0 / 3.0) * H
    dg = torch.clamp(f_tr, min=0.0) / denom
    return 0.5 * bulk * tr * tr + mu * nrm * nrm - 0.5 * denom * dg * dg


def j2_grad_u(points, conn, etype, qorder, u):
    """graddisplacement [E,Q,D,D] (element_assembler.py:631-633): row = component, column = derivative."""
    _, grad, jxw, _ = shape_grad_jxw(points, conn, etype, qorder)
    return torch.einsum("eba,eqbd->eqad", u[conn], grad), grad, jxw


def j2_energy(points, conn, etype, qorder, u, eps_p, alpha, E=200e9, nu=0.3, sig0=250e6, H=1e9):
    """``J2Plasticity.energy`` = sum(psi * jxw) over elements and points
    (element_assembler.py:641-657 with builtin.py:464-511)."""
    G, _, jxw = j2_grad_u(points, conn, etype, qorder, u)
    return (j2_psi(G, eps_p, alpha, E, nu, sig0, H) * jxw).sum()


def j2_element_derivatives(points, conn, etype, qorder, u, eps_p, alpha, E=200e9, nu=0.3, sig0=250e6, H=1e9):
    """Per-element residual [E,B,D] and tangent [E,B,B,D,D] by autograd of the
    element energy -- what ``energy().backward()`` (and a Newton solve's
    autograd Jacobian) produce (element_assembler.py:447-659).  Autograd's
    Hessian is 0 or NaN where the trial deviator vanishes (torch.norm's
    derivative at 0), so callers keep away from that point."""
    _, grad, jxw, _ = shape_grad_jxw(points, conn, etype, qorder)

    def e_energy(ue_, g_, w_, ep_, al_):
        G = torch.einsum("ba,qbd->qad", ue_, g_)
        return (j2_psi(G, ep_, al_, E, nu, sig0, H) * w_).sum()

    ue = u[conn]
    res = torch.func.vmap(torch.func.grad(e_energy))(ue, grad, jxw, eps_p, alpha)
    hes = torch.func.vmap(torch.func.hessian(e_energy))(ue, grad, jxw, eps_p, alpha)
    return res, hes.permute(0, 1, 3, 2, 4).contiguous()


def j2_update(points, conn, etype, qorder, u, eps_p, alpha, E=200e9, nu=0.3, sig0=250e6, H=1e9):
    """Radial-return history update of ``J2Plasticity.update_state``
    (builtin.py:513-561): returns new (eps_p [E,Q,3,3], alpha [E,Q])."""
    mu, _ = j2_moduli(E, nu)
    G, _, _ = j2_grad_u(points, conn, etype, qorder, u)
    eps_tr = _j2_strain3(G) - eps_p
    tr = eps_tr.diagonal(dim1=-2, dim2=-1).sum(-1)
    dev = eps_tr - (tr / 3.0)[..., None, None] * torch.eye(3, dtype=u.dtype)
    nrm = torch.linalg.norm(dev, dim=(-2, -1))
    f_tr = 2.0 * mu * nrm - np.sqrt(2.0 / 3.0) * (sig0 + H * alpha)
    dg = torch.clamp(f_tr, min=0.0) / (2.0 * mu + (2.0 / 3.0) * H)
    safe = torch.where(nrm < 1e-12, torch.ones_like(nrm), nrm)
    dgm = dg * (f_tr > 0).to(dg.dtype)
    return eps_p + dgm[..., None, None] * dev / safe[..., None, None], alpha + np.sqrt(2.0 / 3.0) * dgm


# ------------------------------------------------------- penalty contact ---
def contact_penalty_psi(x: torch.Tensor, kappa: float, normal, offset: float) -> torch.Tensor:
    """kappa/2 <-(n . x - offset)>_+^2 at points x[..., D]: the penalty density
    of the ContactAssembler example (builtin.py:588-597) for a plane n . x = offset."""
    n = torch.as_tensor(normal, dtype=x.dtype)
    pen = torch.clamp(-((x * n).sum(-1) - offset), min=0.0)
    return 0.5 * kappa * pen * pen


def contact_energy(points_ref, x_nodes, conn, boundary_mask, etype, order, kappa, normal, offset):
    """``ContactAssembler.energy`` (builtin.py:609-728) for the penalty
    density: the nodal positions x are interpolated at the facet points of
    the selected facets with the cell basis (einsum 'sbd,sqb->sqd'), and the
    density is integrated against FxW of the configuration points_ref."""
    elem, fac = facet_select(conn.numpy(), boundary_mask.numpy(), etype)
    phi, _, fxw = facet_geometry(points_ref, conn, elem, fac, etype, order)
    dofs = conn[torch.from_numpy(elem)]
    xq = torch.einsum("sbd,sqb->sqd", x_nodes[dofs], phi)
    return (contact_penalty_psi(xq, kappa, normal, offset) * fxw).sum()


def contact_facet_derivatives(points_ref, x_nodes, conn, boundary_mask, etype, order, kappa, normal, offset):
    """Per-facet gradient [F, B, D] and Hessian [F, B, B, D, D] of the penalty
    contact energy with respect to the nodal positions, by autograd (what
    ``energy().backward()`` and an autograd Jacobian give), plus the facets'
    (element, local facet) lists."""
    elem, fac = facet_select(conn.numpy(), boundary_mask.numpy(), etype)
    phi, _, fxw = facet_geometry(points_ref, conn, elem, fac, etype, order)
    dofs = conn[torch.from_numpy(elem)]

    def f_energy(xe, ph, w):
        xq = torch.einsum("bd,qb->qd", xe, ph)
        return (contact_penalty_psi(xq, kappa, normal, offset) * w).sum()

    xe = x_nodes[dofs]
    res = torch.func.vmap(torch.func.grad(f_energy))(xe, phi, fxw)
    hes = torch.func.vmap(torch.func.hessian(f_energy))(xe, phi, fxw)
    return res, hes.permute(0, 1, 3, 2, 4).contiguous(), elem, fac
