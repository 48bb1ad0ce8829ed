"""Reference-element data for the simplices on the assembly path.

* Quadrature rules on the reference triangle / tetrahedron, restated from
  the reference tables (tensormesh/element/quadrature.py:91-160) for rule
  orders 1-4.  ``quadrature_order`` is a rule *order*, not a point count:
  triangle 1/2/3/4 -> 1/3/4/6 points, tetrahedron 1/2/3/4 -> 1/4/5/11 points;
  the order-3 rules carry a negative weight.
* Lagrange shape functions in closed form for P1 and P2, in the reference's
  internal node order (vertices, then edge midpoints in the order of
  ``Triangle.edge = [[1,2],[0,2],[0,1]]`` / ``Tetrahedron.edge =
  [[2,3],[1,3],[1,2],[0,3],[0,2],[0,1]]``, tensormesh/element/element.py:1684
  and :1949).  The reference obtains the same polynomials by inverting a
  Vandermonde matrix (tensormesh/element/element.py:555-659).

``reference_tables`` packs weights, values and reference gradients into the
flat float64 layout that the HIP kernels read (documented in
``include/tensormesh_amd.h``).

Facets (boundary integrals, tensormesh/assemble/facet_assembler.py):

* triangle facets are the edges ``[[1,2],[0,2],[0,1]]`` and tetrahedron
  facets the faces ``((1,2,3),(0,2,3),(0,1,3),(0,1,2))``
  (tensormesh/element/element.py:1684-1685, :1949-1950), each with the
  element's edge-midpoint nodes on it for P2;
* a facet is the affine image gamma_f(r) = P_a + sum_k r_k (P_{k+1} - P_a) of the
  reference facet (reference-element vertices P, element.py:972-1067);
* triangle facets use an n-point Gauss-Legendre rule on [0, 1] with n equal
  to ``quadrature_order`` (1-7; quadrature.py:13-66), tetrahedron facets the
  triangle rule of that order (quadrature.py:362-376).
"""

from __future__ import annotations

import functools
from typing import Tuple

import numpy as np
import torch

TRI_QUAD = {
    1: ([0.5], [[0.3333333333333333, 0.3333333333333333]]),
    2: ([0.16666666666666666] * 3,
        [[0.16666666666666666, 0.16666666666666666],
         [0.6666666666666666, 0.16666666666666666],
         [0.16666666666666666, 0.6666666666666666]]),
    3: ([-0.28125, 0.260416666666667, 0.260416666666667, 0.260416666666667],
        [[0.333333333333333, 0.333333333333333], [0.2, 0.6], [0.6, 0.2], [0.2, 0.2]]),
    4: ([0.111690794839006, 0.054975871827661, 0.111690794839006,
         0.054975871827661, 0.111690794839006, 0.054975871827661],
        [[0.445948490915965, 0.10810301816807], [0.0915762135097699, 0.816847572980459],
         [0.10810301816807, 0.445948490915965], [0.816847572980459, 0.091576213509771],
         [0.445948490915965, 0.445948490915965], [0.091576213509771, 0.0915762135097699]]),
    5: ([0.1125, 0.066197076394253, 0.0629695902724135, 0.066197076394253, 0.0629695902724135,
         0.066197076394253, 0.0629695902724135],
        [[0.333333333333333, 0.333333333333333], [0.470142064105115, 0.05971587178977],
         [0.101286507323457, 0.797426985353087], [0.05971587178977, 0.470142064105115],
         [0.797426985353087, 0.101286507323456], [0.470142064105115, 0.470142064105115],
         [0.101286507323456, 0.101286507323457]]),
    6: ([0.0583931378631895, 0.0254224531851035, 0.0583931378631895, 0.0254224531851035, 0.0583931378631895,
         0.0254224531851035] + [0.041425537809187] * 6,
        [[0.249286745170911, 0.501426509658179], [0.0630890144915021, 0.873821971016996],
         [0.501426509658179, 0.24928674517091], [0.873821971016996, 0.063089014491502],
         [0.24928674517091, 0.249286745170911], [0.063089014491502, 0.0630890144915021],
         [0.636502499121399, 0.053145049844817], [0.310352451033784, 0.053145049844817],
         [0.053145049844817, 0.310352451033784], [0.053145049844817, 0.636502499121399],
         [0.310352451033784, 0.636502499121399], [0.636502499121399, 0.310352451033784]]),
    7: ([-0.074785022233841] + [0.087807628716604, 0.026673617804419] * 3 + [0.0385568804451285] * 6,
        [[0.333333333333333, 0.333333333333333], [0.26034596607904, 0.47930806784192],
         [0.065130102902216, 0.869739794195568], [0.47930806784192, 0.26034596607904],
         [0.869739794195568, 0.065130102902216], [0.26034596607904, 0.26034596607904],
         [0.065130102902216, 0.065130102902216], [0.63844418856981, 0.048690315425316],
         [0.312865496004874, 0.048690315425316], [0.048690315425316, 0.312865496004874],
         [0.048690315425316, 0.63844418856981], [0.312865496004874, 0.63844418856981],
         [0.63844418856981, 0.312865496004874]]),
}

TET_QUAD = {
    1: ([0.16666666666666666], [[0.25, 0.25, 0.25]]),
    2: ([0.041666666666666664] * 4,
        [[0.5854101966249685, 0.1381966011250105, 0.1381966011250105],
         [0.1381966011250105, 0.1381966011250105, 0.1381966011250105],
         [0.1381966011250105, 0.1381966011250105, 0.5854101966249685],
         [0.1381966011250105, 0.5854101966249685, 0.1381966011250105]]),
    3: ([-0.13333333333333333, 0.075, 0.075, 0.075, 0.075],
        [[0.25, 0.25, 0.25], [0.5, 0.1666666666666667, 0.1666666666666667],
         [0.1666666666666667, 0.1666666666666667, 0.1666666666666667],
         [0.1666666666666667, 0.1666666666666667, 0.5],
         [0.1666666666666667, 0.5, 0.1666666666666667]]),
    4: ([-0.01315555555555555] + [0.007622222222222217] * 4 + [0.02488888888888888] * 6,
        [[0.25, 0.25, 0.25],
         [0.7857142857142857, 0.0714285714285714, 0.0714285714285714],
         [0.0714285714285714, 0.0714285714285714, 0.0714285714285714],
         [0.0714285714285714, 0.0714285714285714, 0.7857142857142857],
         [0.0714285714285714, 0.7857142857142857, 0.0714285714285714],
         [0.1005964238332008, 0.3994035761667992, 0.3994035761667992],
         [0.3994035761667992, 0.1005964238332008, 0.3994035761667992],
         [0.3994035761667992, 0.3994035761667992, 0.1005964238332008],
         [0.3994035761667992, 0.1005964238332008, 0.1005964238332008],
         [0.1005964238332008, 0.3994035761667992, 0.1005964238332008],
         [0.1005964238332008, 0.1005964238332008, 0.3994035761667992]]),
    5: ([0.003174603174603167] * 6 + [0.014764970790496783] * 4 + [0.022139791114265117] * 4,
        [[0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0], [0.5, 0.0, 0.0], [0.0, 0.5, 0.0], [0.0, 0.0, 0.5],
         [0.6984197043243866, 0.1005267652252045, 0.1005267652252045],
         [0.1005267652252045, 0.1005267652252045, 0.1005267652252045],
         [0.1005267652252045, 0.1005267652252045, 0.6984197043243866],
         [0.1005267652252045, 0.6984197043243866, 0.1005267652252045],
         [0.0568813795204234, 0.3143728734931922, 0.3143728734931922],
         [0.3143728734931922, 0.3143728734931922, 0.3143728734931922],
         [0.3143728734931922, 0.3143728734931922, 0.0568813795204234],
         [0.3143728734931922, 0.0568813795204234, 0.3143728734931922]]),
    6: ([0.030283678097089182] + [0.006026785714285717] * 4 + [0.011645249086028967] * 4
        + [0.010949141561386449] * 6,
        [[0.25, 0.25, 0.25], [0.0, 0.3333333333333333, 0.3333333333333333],
         [0.3333333333333333, 0.3333333333333333, 0.3333333333333333],
         [0.3333333333333333, 0.3333333333333333, 0.0], [0.3333333333333333, 0.0, 0.3333333333333333],
         [0.7272727272727273, 0.0909090909090909, 0.0909090909090909],
         [0.0909090909090909, 0.0909090909090909, 0.0909090909090909],
         [0.0909090909090909, 0.0909090909090909, 0.7272727272727273],
         [0.0909090909090909, 0.7272727272727273, 0.0909090909090909],
         [0.4334498464263357, 0.0665501535736643, 0.0665501535736643],
         [0.0665501535736643, 0.4334498464263357, 0.0665501535736643],
         [0.0665501535736643, 0.0665501535736643, 0.4334498464263357],
         [0.0665501535736643, 0.4334498464263357, 0.4334498464263357],
         [0.4334498464263357, 0.0665501535736643, 0.4334498464263357],
         [0.4334498464263357, 0.4334498464263357, 0.0665501535736643]]),
    7: ([0.0066537917096946494] * 4 + [0.0016795351758867834] * 4 + [0.009226196923942399] * 4
        + [0.008035714285714283] * 12,
        [[0.3561913862225449, 0.2146028712591517, 0.2146028712591517],
         [0.2146028712591517, 0.2146028712591517, 0.2146028712591517],
         [0.2146028712591517, 0.2146028712591517, 0.3561913862225449],
         [0.2146028712591517, 0.3561913862225449, 0.2146028712591517],
         [0.877978124396166, 0.0406739585346113, 0.0406739585346113],
         [0.0406739585346113, 0.0406739585346113, 0.0406739585346113],
         [0.0406739585346113, 0.0406739585346113, 0.877978124396166],
         [0.0406739585346113, 0.877978124396166, 0.0406739585346113],
         [0.0329863295731731, 0.3223378901422757, 0.3223378901422757],
         [0.3223378901422757, 0.3223378901422757, 0.3223378901422757],
         [0.3223378901422757, 0.3223378901422757, 0.0329863295731731],
         [0.3223378901422757, 0.0329863295731731, 0.3223378901422757],
         [0.2696723314583159, 0.0636610018750175, 0.0636610018750175],
         [0.0636610018750175, 0.2696723314583159, 0.0636610018750175],
         [0.0636610018750175, 0.0636610018750175, 0.2696723314583159],
         [0.6030056647916491, 0.0636610018750175, 0.0636610018750175],
         [0.0636610018750175, 0.6030056647916491, 0.0636610018750175],
         [0.0636610018750175, 0.0636610018750175, 0.6030056647916491],
         [0.0636610018750175, 0.2696723314583159, 0.6030056647916491],
         [0.2696723314583159, 0.6030056647916491, 0.0636610018750175],
         [0.6030056647916491, 0.0636610018750175, 0.2696723314583159],
         [0.0636610018750175, 0.6030056647916491, 0.2696723314583159],
         [0.2696723314583159, 0.0636610018750175, 0.6030056647916491],
         [0.6030056647916491, 0.2696723314583159, 0.0636610018750175]]),
}

# n-point Gauss-Legendre on [-1, 1] (quadrature.py:36-54), mapped to [0, 1] with
# weights halved (quadrature.py:59-61)
_GL = {
    1: ([2.0], [0.0]),
    2: ([1.0, 1.0], [-0.57735026918962576451, 0.57735026918962576451]),
    3: ([0.55555555555555555556, 0.88888888888888888889, 0.55555555555555555556],
        [-0.77459666924148337704, 0.0, 0.77459666924148337704]),
    4: ([0.34785484513745385737, 0.65214515486254614263, 0.65214515486254614263, 0.34785484513745385737],
        [-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526]),
    5: ([0.23692688505618908751, 0.47862867049936646804, 0.56888888888888888889, 0.47862867049936646804,
         0.23692688505618908751],
        [-0.9061798459386640, -0.5384693101056831, 0.0, 0.5384693101056831, 0.9061798459386640]),
    6: ([0.17132449237917034504, 0.36076157304813860757, 0.46791393457269104739, 0.46791393457269104739,
         0.36076157304813860757, 0.17132449237917034504],
        [-0.9324695142031521, -0.6612093864662645, -0.2386191860831969, 0.2386191860831969, 0.6612093864662645,
         0.9324695142031521]),
    7: ([0.12948496616886969327, 0.27970539148927666790, 0.38183005050511894495, 0.41795918367346938776,
         0.38183005050511894495, 0.27970539148927666790, 0.12948496616886969327],
        [-0.9491079123427585, -0.7415311855993945, -0.4058451513773972, 0.0, 0.4058451513773972,
         0.7415311855993945, 0.9491079123427585]),
}

ELEMENT_SHAPE = {
    # element type: (dim, order, n_basis, family)
    "triangle": (2, 1, 3, "tri"),
    "triangle6": (2, 2, 6, "tri"),
    "tetra": (3, 1, 4, "tet"),
    "tetra10": (3, 2, 10, "tet"),
    "triangle10": (2, 3, 10, "tri"),
    "tetra20": (3, 3, 20, "tet"),
    "quad": (2, 1, 4, "quad"),
    "quad9": (2, 2, 9, "quad"),
    "hexahedron": (3, 1, 8, "hex"),
    "hexahedron27": (3, 2, 27, "hex"),
}
TENSOR_FAMILIES = ("quad", "hex")


def _tensor_rule(dim: int, order: int) -> Tuple[np.ndarray, np.ndarray]:
    """Tensor product of the ``order``-point Gauss-Legendre rule on [0, 1] in
    the point order of quad_quadrature / hex_quadrature (quadrature.py:163-239,
    torch.meshgrid(..., indexing='xy')): quadrilateral x fastest, then y;
    hexahedron z fastest, then x, then y."""
    w1, p1 = line_quadrature(order)
    p1 = p1[:, 0]
    n = p1.size
    if dim == 2:
        iy, ix = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        iy, ix = iy.ravel(), ix.ravel()
        return w1[ix] * w1[iy], np.stack([p1[ix], p1[iy]], 1)
    iy, ix, iz = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    iy, ix, iz = iy.ravel(), ix.ravel(), iz.ravel()
    return w1[ix] * w1[iy] * w1[iz], np.stack([p1[ix], p1[iy], p1[iz]], 1)


def quadrature(element_type: str, order: int) -> Tuple[np.ndarray, np.ndarray]:
    """(weights[Q], points[Q, D]) as float64 numpy arrays.  For quadrilaterals
    and hexahedra ``order`` is the number of Gauss points per axis (1-7)."""
    dim, _, _, fam = ELEMENT_SHAPE[element_type]
    if fam in TENSOR_FAMILIES:
        if order not in _GL:
            raise NotImplementedError(f"quadrature_order={order} for {element_type}: supported orders are 1-7")
        return _tensor_rule(dim, order)
    table = TRI_QUAD if fam == "tri" else TET_QUAD
    if order not in table:
        raise NotImplementedError(
            f"quadrature_order={order} for {element_type}: supported orders are {sorted(table)}")
    w, p = table[order]
    return np.asarray(w, dtype=np.float64), np.asarray(p, dtype=np.float64)


# Tensor-product cells on [0, 1]^D (element.py:1811-1814, 2100-2104): vertices in
# lexicographic order (x fastest), edges and faces as the reference lists them.
_QUAD_EDGES = ((0, 1), (0, 2), (1, 3), (2, 3))
_HEX_EDGES = ((0, 1), (0, 2), (0, 4), (1, 3), (1, 5), (2, 3), (2, 6), (3, 7), (4, 5), (4, 6), (5, 7), (6, 7))
_HEX_FACES = ((0, 1, 2, 3), (0, 1, 4, 5), (0, 2, 4, 6), (1, 3, 5, 7), (2, 3, 6, 7), (4, 5, 6, 7))


@functools.lru_cache(maxsize=None)
def tensor_nodes(element_type: str) -> np.ndarray:
    """Lagrange nodes [B, D] of a tensor-product cell in the reference's basis
    order (basis.py:697-748, 861-905): vertices, edge nodes (edge order), face
    nodes (face order), cell node; for order 2 these are midpoints/centres."""
    dim, order, nb, fam = ELEMENT_SHAPE[element_type]
    verts = np.array([[(v >> d) & 1 for d in range(dim)] for v in range(2 ** dim)], np.float64)
    if order == 1:
        return verts
    edges = _QUAD_EDGES if fam == "quad" else _HEX_EDGES
    nodes = [verts] + [0.5 * (verts[a] + verts[b])[None] for a, b in edges]
    faces = ((0, 1, 2, 3),) if fam == "quad" else _HEX_FACES
    nodes += [verts[list(f)].mean(0)[None] for f in faces]
    if fam == "hex":
        nodes.append(np.full((1, 3), 0.5))
    return np.concatenate(nodes, 0)


def _lagrange_1d(t: np.ndarray, order: int):
    """Values and derivatives [len(t), order+1] of the 1-D Lagrange polynomials
    on the nodes k / order."""
    nodes = np.arange(order + 1) / order
    val = np.ones((t.size, order + 1))
    der = np.zeros((t.size, order + 1))
    for k in range(order + 1):
        others = [m for m in range(order + 1) if m != k]
        for m in others:
            val[:, k] *= (t - nodes[m]) / (nodes[k] - nodes[m])
        for m in others:
            term = np.ones(t.size) / (nodes[k] - nodes[m])
            for m2 in others:
                if m2 != m:
                    term = term * (t - nodes[m2]) / (nodes[k] - nodes[m2])
            der[:, k] += term
    return val, der


def _tensor_shape(element_type: str, points: np.ndarray):
    """phi[Q, B], dphi[Q, D, B] of a tensor-product cell: node n at grid index
    (k_0, .., k_{D-1}) has phi_n = prod_d L_{k_d}(xi_d)."""
    dim, order, nb, _ = ELEMENT_SHAPE[element_type]
    idx = np.rint(tensor_nodes(element_type) * order).astype(np.int64)      # [B, D]
    vals, ders = zip(*[_lagrange_1d(points[:, d], order) for d in range(dim)])
    phi = np.ones((points.shape[0], nb))
    for d in range(dim):
        phi *= vals[d][:, idx[:, d]]
    dphi = np.ones((points.shape[0], dim, nb))
    for g in range(dim):
        for d in range(dim):
            dphi[:, g] *= (ders[d] if d == g else vals[d])[:, idx[:, d]]
    return phi, dphi


def _barycentric(points: np.ndarray) -> np.ndarray:
    return np.concatenate([1.0 - points.sum(1, keepdims=True), points], 1)  # [Q, D+1]


def shape_values(element_type: str, points: np.ndarray) -> np.ndarray:
    """phi[Q, B] at reference points."""
    dim, order, nb, fam = ELEMENT_SHAPE[element_type]
    if fam in TENSOR_FAMILIES:
        return _tensor_shape(element_type, points)[0]
    lam = _barycentric(points)
    if order == 1:
        return lam.copy()
    if order == 3:
        return _p3_shape(fam, lam, None)[0]
    edges = ((1, 2), (0, 2), (0, 1)) if fam == "tri" else ((2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1))
    vert = lam * (2.0 * lam - 1.0)
    mid = np.stack([4.0 * lam[:, a] * lam[:, b] for a, b in edges], 1)
    return np.concatenate([vert, mid], 1)


def shape_ref_grads(element_type: str, points: np.ndarray) -> np.ndarray:
    """dphi/dxi [Q, D, B] at reference points (layout of the reference's
    ``Polynomials.map`` output, tensormesh/element/element.py:960)."""
    dim, order, nb, fam = ELEMENT_SHAPE[element_type]
    if fam in TENSOR_FAMILIES:
        return _tensor_shape(element_type, points)[1]
    q = points.shape[0]
    lam = _barycentric(points)
    # d lambda_k / d xi_i : lambda_0 = 1 - sum(xi), lambda_k = xi_{k-1}
    dlam = np.zeros((dim, dim + 1))
    dlam[:, 0] = -1.0
    for i in range(dim):
        dlam[i, i + 1] = 1.0
    out = np.zeros((q, dim, nb))
    if order == 1:
        out[:] = dlam[None]
        return out
    if order == 3:
        return _p3_shape(fam, lam, dlam)[1]
    edges = ((1, 2), (0, 2), (0, 1)) if fam == "tri" else ((2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1))
    for k in range(dim + 1):
        out[:, :, k] = (4.0 * lam[:, k:k + 1] - 1.0) * dlam[None, :, k]
    for m, (a, b) in enumerate(edges):
        out[:, :, dim + 1 + m] = 4.0 * (lam[:, b:b + 1] * dlam[None, :, a] + lam[:, a:a + 1] * dlam[None, :, b])
    return out


_SIMPLEX_FACES = {"tri": ((0, 1, 2),), "tet": ((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2))}


def _p3_shape(fam: str, lam: np.ndarray, dlam):
    """Cubic Lagrange basis in barycentric coordinates, in the reference's node
    order (tri_basis / tet_basis, basis.py:639-859): vertices
    l_i (3 l_i - 1)(3 l_i - 2) / 2; two nodes per edge (a, b) at 1/3 and 2/3
    from a, 9/2 l_a l_b (3 l_a - 1) and 9/2 l_a l_b (3 l_b - 1); one node per
    face (i, j, k) at its centroid, 27 l_i l_j l_k.  Returns (phi [Q, B],
    dphi [Q, D, B] when dlam [D, D+1] is given)."""
    edges = ((1, 2), (0, 2), (0, 1)) if fam == "tri" else ((2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1))
    terms = []   # each basis function as a product of (coefficient, [(k, a, b)]) factors a*l_k + b
    for i in range(lam.shape[1]):
        terms.append((0.5, [(i, 1.0, 0.0), (i, 3.0, -1.0), (i, 3.0, -2.0)]))
    for a, b in edges:
        terms.append((4.5, [(a, 1.0, 0.0), (b, 1.0, 0.0), (a, 3.0, -1.0)]))
        terms.append((4.5, [(a, 1.0, 0.0), (b, 1.0, 0.0), (b, 3.0, -1.0)]))
    for i, j, k in _SIMPLEX_FACES[fam]:
        terms.append((27.0, [(i, 1.0, 0.0), (j, 1.0, 0.0), (k, 1.0, 0.0)]))
    q = lam.shape[0]
    phi = np.zeros((q, len(terms)))
    dphi = None if dlam is None else np.zeros((q, dlam.shape[0], len(terms)))
    for n, (c, fac) in enumerate(terms):
        vals = [a * lam[:, k] + b for k, a, b in fac]
        phi[:, n] = c * np.prod(vals, axis=0)
        if dlam is not None:
            for m, (k, a, b) in enumerate(fac):
                others = np.prod([v for t, v in enumerate(vals) if t != m], axis=0)
                dphi[:, :, n] += c * a * others[:, None] * dlam[None, :, k]
    return phi, dphi


@functools.lru_cache(maxsize=None)
def reference_tables(element_type: str, qorder: int) -> np.ndarray:
    """Flat float64 table read by the kernels.

    Layout (Q quadrature points, B basis functions, D dims)::

        [0, Q)                 w[q]
        [Q, Q + Q*B)           phi[q][b]
        [Q + Q*B, Q + Q*B + Q*D*B)   dphi[q][d][b]   (d/dxi_d)
    """
    w, p = quadrature(element_type, qorder)
    phi = shape_values(element_type, p)
    dphi = shape_ref_grads(element_type, p)
    return np.concatenate([w, phi.ravel(), dphi.ravel()]).astype(np.float64)


def mass_reference(element_type: str, qorder: int) -> np.ndarray:
    """Mref[B, B] = sum_q w_q phi_i(q) phi_j(q) on the reference element."""
    w, p = quadrature(element_type, qorder)
    phi = shape_values(element_type, p)
    return np.einsum("q,qi,qj->ij", w, phi, phi)


def torch_tables(element_type: str, qorder: int, dtype, device):
    """(w[Q], phi[Q,B], dphi[Q,D,B]) as torch tensors for the torch form path."""
    w, p = quadrature(element_type, qorder)
    return (torch.tensor(w, dtype=dtype, device=device),
            torch.tensor(shape_values(element_type, p), dtype=dtype, device=device),
            torch.tensor(shape_ref_grads(element_type, p), dtype=dtype, device=device))


def line_quadrature(order: int) -> Tuple[np.ndarray, np.ndarray]:
    """(weights[Q], points[Q, 1]) of the n-point Gauss-Legendre rule on [0, 1]."""
    if order not in _GL:
        raise NotImplementedError(f"facet quadrature_order={order} on lines: supported orders are 1-7")
    w, p = _GL[order]
    return np.asarray(w, np.float64) / 2.0, ((np.asarray(p, np.float64) + 1.0) / 2.0)[:, None]


_REF_VERTS = {"tri": np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]]),
              "tet": np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]),
              "quad": np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [1.0, 1.0]]),
              "hex": np.array([[(v >> d) & 1 for d in range(3)] for v in range(8)], np.float64)}
# facets: triangle / quadrilateral edges, tetrahedron / hexahedron faces
_FACET_VERTS = {"tri": ((1, 2), (0, 2), (0, 1)), "tet": ((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)),
                "quad": _QUAD_EDGES, "hex": _HEX_FACES}
_EDGES = {"tri": ((1, 2), (0, 2), (0, 1)), "tet": ((2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1)),
          "quad": _QUAD_EDGES, "hex": _HEX_EDGES}


def facet_nodes(element_type: str) -> np.ndarray:
    """[n_facet, n_node_per_facet] local nodes of each facet: its vertices,
    then (order 2) the midpoint nodes of the element edges lying on it and, for
    hexahedron faces, the face's centre node (basis.py:1127-1287)."""
    dim, order, nb, fam = ELEMENT_SHAPE[element_type]
    n_vert = _REF_VERTS[fam].shape[0]
    out = []
    for fi, fv in enumerate(_FACET_VERTS[fam]):
        nodes = list(fv)
        if order == 2:
            nodes += [n_vert + m for m, (a, b) in enumerate(_EDGES[fam]) if a in fv and b in fv]
            if fam == "hex":
                nodes.append(n_vert + len(_HEX_EDGES) + fi)
        elif order == 3:
            for m, (a, b) in enumerate(_EDGES[fam]):
                if a in fv and b in fv:
                    nodes += [n_vert + 2 * m, n_vert + 2 * m + 1]
            if fam == "tet":
                nodes.append(n_vert + 2 * len(_EDGES[fam]) + fi)
        out.append(nodes)
    return np.asarray(out, np.int64)


@functools.lru_cache(maxsize=None)
def facet_tables(element_type: str, qorder: int):
    """Facet quadrature data in cell reference coordinates:
    (w[Qf], phi[nF, Qf, B], dphi[nF, Qf, D, B], dgamma[nF, D-1, D]) with
    dgamma[f][k] = d gamma_f / d r_k (the facet's reference tangents)."""
    dim, order, nb, fam = ELEMENT_SHAPE[element_type]
    if fam in ("tri", "quad"):
        w, r = line_quadrature(qorder)                    # facet_quadrature_2d
    elif fam == "tet":
        w, r = quadrature("triangle", qorder)             # tet_facet_quadrature
    else:
        w, r = quadrature("quad", qorder)                 # hex_facet_quadrature
    P = _REF_VERTS[fam]
    phi, dphi, dgam = [], [], []
    for fv in _FACET_VERTS[fam]:
        a = P[fv[0]]
        tang = np.stack([P[fv[k + 1]] - a for k in range(dim - 1)])   # [D-1, D]
        xi = a[None, :] + r @ tang                                     # [Qf, D]
        phi.append(shape_values(element_type, xi))
        dphi.append(shape_ref_grads(element_type, xi))
        dgam.append(tang)
    return (np.asarray(w, np.float64), np.stack(phi), np.stack(dphi), np.stack(dgam))


@functools.lru_cache(maxsize=None)
def facet_reference_tables(element_type: str, qorder: int) -> np.ndarray:
    """Flat float64 facet table read by the facet kernel::

        [0, Qf)                                w[q]
        then nF*Qf*B                           phi[f][q][b]
        then nF*Qf*D*B                         dphi[f][q][d][b]
        then nF*(D-1)*D                        dgamma[f][k][d]
    """
    w, phi, dphi, dgam = facet_tables(element_type, qorder)
    return np.concatenate([w, phi.ravel(), dphi.ravel(), dgam.ravel()]).astype(np.float64)
