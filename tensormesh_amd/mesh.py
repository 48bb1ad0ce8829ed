"""Mesh container and structured simplex generators.

The reference's ``Mesh`` (tensormesh/mesh/mesh.py:25-147) wraps meshio and
builds its meshes with gmsh (tensormesh/dataset/mesh/rectangle.py:49-111).
Neither library is part of this stack, so this module keeps the part of the
``Mesh`` interface that the assembly path reads -- ``points``, ``cells``,
``elements()``, ``default_element_type``, ``boundary_mask``, ``n_points``,
``dtype``, ``device``, ``.to()`` -- and generates structured simplex meshes
directly:

* ``unit_square_tri``: an ``n x n`` grid of squares, each split into the two
  triangles ``(a, b, c)`` and ``(a, c, d)`` with ``a=(i,j)``, ``b=(i+1,j)``,
  ``c=(i+1,j+1)``, ``d=(i,j+1)``; nodes numbered x-fastest.
* ``unit_cube_tet``: an ``n^3`` grid of cubes, each split into the six Kuhn
  tetrahedra along the main diagonal (one per permutation of the axes);
  odd permutations have two vertices swapped so every tetrahedron is
  positively oriented.
* ``to_quadratic``: adds one node per unique edge and returns ``triangle6`` /
  ``tetra10`` connectivity in the reference's internal node order (vertices,
  then edge midpoints in the order of ``Triangle.edge`` /
  ``Tetrahedron.edge``, tensormesh/element/element.py:1684 and :1949).

Interior nodes can be jittered uniformly by ``+-jitter*h`` per coordinate
with a seeded numpy generator, which is how the test and benchmark
workloads are defined.
"""

from __future__ import annotations

import itertools
from typing import Dict, Optional

import numpy as np
import torch

# Reference-element edge tables in the reference's internal order.
TRI_EDGES = ((1, 2), (0, 2), (0, 1))
TET_EDGES = ((2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1))

ELEMENT_INFO = {
    # element type: (spatial dim, basis order, nodes per element)
    "triangle": (2, 1, 3),
    "tetra": (3, 1, 4),
    "triangle6": (2, 2, 6),
    "tetra10": (3, 2, 10),
    "triangle10": (2, 3, 10),
    "tetra20": (3, 3, 20),
    "quad": (2, 1, 4),
    "quad9": (2, 2, 9),
    "hexahedron": (3, 1, 8),
    "hexahedron27": (3, 2, 27),
}


class Mesh:
    """Points plus one or more cell blocks.

    Parameters
    ----------
    points : torch.Tensor
        ``[N, D]`` node coordinates.
    cells : dict[str, torch.Tensor]
        element type -> ``[E, B]`` connectivity (any integer dtype).
    boundary_mask : torch.Tensor, optional
        ``[N]`` bool, True on Dirichlet boundary nodes.
    """

    def __init__(self, points: torch.Tensor, cells: Dict[str, torch.Tensor],
                 boundary_mask: Optional[torch.Tensor] = None):
        assert points.dim() == 2, f"points must be [N, D], got {tuple(points.shape)}"
        for k, v in cells.items():
            assert k in ELEMENT_INFO, f"unsupported element type {k}"
            assert v.dim() == 2 and v.shape[1] == ELEMENT_INFO[k][2], \
                f"{k} connectivity must be [E, {ELEMENT_INFO[k][2]}], got {tuple(v.shape)}"
            assert ELEMENT_INFO[k][0] == points.shape[1], \
                f"{k} needs {ELEMENT_INFO[k][0]}-D points, got {points.shape[1]}-D"
        self.points = points
        self.cells = {k: v.long() for k, v in cells.items()}
        if boundary_mask is None:
            boundary_mask = torch.zeros(points.shape[0], dtype=torch.bool, device=points.device)
        self.boundary_mask = boundary_mask

    # --- reference-compatible accessors (tensormesh/mesh/mesh.py:403-486) ---
    @property
    def default_element_type(self) -> str:
        return next(iter(self.cells.keys()))

    def elements(self, element_type: Optional[str] = None):
        if element_type is not None:
            return self.cells[element_type]
        if len(self.cells) == 1:
            return self.cells[self.default_element_type]
        return dict(self.cells)

    @property
    def n_points(self) -> int:
        return int(self.points.shape[0])

    @property
    def dim(self) -> int:
        return int(self.points.shape[1])

    @property
    def dtype(self) -> torch.dtype:
        return self.points.dtype

    @property
    def device(self) -> torch.device:
        return self.points.device

    def to(self, *args, **kwargs) -> "Mesh":
        pts = self.points.to(*args, **kwargs)
        dev = pts.device
        return Mesh(pts, {k: v.to(dev) for k, v in self.cells.items()},
                    self.boundary_mask.to(dev))

    def cuda(self) -> "Mesh":
        return self.to("cuda")

    def double(self) -> "Mesh":
        return self.to(torch.float64)

    def float(self) -> "Mesh":
        return self.to(torch.float32)

    def __repr__(self) -> str:
        blocks = ", ".join(f"{k}:{tuple(v.shape)}" for k, v in self.cells.items())
        return f"Mesh(n_points={self.n_points}, dim={self.dim}, cells={{{blocks}}}, dtype={self.dtype}, device={self.device})"


def _jitter(points: np.ndarray, interior: np.ndarray, amount: float, seed: Optional[int]) -> np.ndarray:
    if amount == 0.0:
        return points
    rng = np.random.default_rng(seed)
    delta = rng.uniform(-amount, amount, size=points.shape)
    delta[~interior] = 0.0
    return points + delta


def unit_square_tri_arrays(n: int, jitter: float = 0.0, seed: Optional[int] = 0):
    """numpy (points[N,2] f64, cells[E,3] i64, boundary[N] bool) of the split grid."""
    h = 1.0 / n
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="xy")
    # x-fastest numbering: node(i, j) = i + (n+1) j
    pts = np.stack([ii.ravel() * h, jj.ravel() * h], axis=1).astype(np.float64)
    boundary = (ii.ravel() == 0) | (ii.ravel() == n) | (jj.ravel() == 0) | (jj.ravel() == n)
    pts = _jitter(pts, ~boundary, jitter * h, seed)
    ci, cj = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    ci, cj = ci.ravel(), cj.ravel()
    a = ci + (n + 1) * cj
    b = a + 1
    c = a + (n + 1) + 1
    d = a + (n + 1)
    cells = np.empty((2 * n * n, 3), dtype=np.int64)
    cells[0::2] = np.stack([a, b, c], 1)
    cells[1::2] = np.stack([a, c, d], 1)
    return pts, cells, boundary


def unit_cube_tet_arrays(n: int, jitter: float = 0.0, seed: Optional[int] = 0):
    """numpy (points[N,3] f64, cells[E,4] i64, boundary[N] bool) of the Kuhn split."""
    h = 1.0 / n
    m = n + 1
    kk, jj, ii = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    pts = np.stack([ii * h, jj * h, kk * h], axis=1).astype(np.float64)
    boundary = (ii == 0) | (ii == n) | (jj == 0) | (jj == n) | (kk == 0) | (kk == n)
    pts = _jitter(pts, ~boundary, jitter * h, seed)

    ck, cj, ci = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    base = (ci + m * cj + m * m * ck).ravel()
    step = (1, m, m * m)
    tets = []
    for perm in itertools.permutations(range(3)):
        v0 = base
        v1 = v0 + step[perm[0]]
        v2 = v1 + step[perm[1]]
        v3 = base + 1 + m + m * m
        # sign of the Kuhn simplex equals the sign of the permutation
        inversions = sum(1 for x in range(3) for y in range(x + 1, 3) if perm[x] > perm[y])
        if inversions % 2:
            tets.append(np.stack([v0, v2, v1, v3], 1))
        else:
            tets.append(np.stack([v0, v1, v2, v3], 1))
    cells = np.stack(tets, axis=1).reshape(-1, 4).astype(np.int64)
    return pts, cells, boundary


def quadratic_arrays(points: np.ndarray, cells: np.ndarray, boundary: np.ndarray):
    """Promote P1 simplices to P2 (triangle6 / tetra10) with midpoint nodes.

    Midpoint node ids are appended after the vertices, numbered in the order
    of first appearance of each unique edge (element-major, local edge order).
    A midpoint is a boundary node iff both of its end vertices are.
    """
    b = cells.shape[1]
    edges = TRI_EDGES if b == 3 else TET_EDGES
    loc = np.asarray(edges, dtype=np.int64)
    e0 = cells[:, loc[:, 0]]
    e1 = cells[:, loc[:, 1]]
    lo = np.minimum(e0, e1).ravel()
    hi = np.maximum(e0, e1).ravel()
    n = points.shape[0]
    key = lo * n + hi
    uniq, first_idx, inv = np.unique(key, return_index=True, return_inverse=True)
    # renumber unique edges by first appearance
    order = np.argsort(first_idx, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    mid_id = n + rank[inv]
    ulo = uniq[order] // n
    uhi = uniq[order] % n
    mids = 0.5 * (points[ulo] + points[uhi])
    new_points = np.concatenate([points, mids], 0)
    new_boundary = np.concatenate([boundary, boundary[ulo] & boundary[uhi]])
    new_cells = np.concatenate([cells, mid_id.reshape(cells.shape[0], len(edges))], 1)
    return new_points, new_cells, new_boundary


def cubic_arrays(points: np.ndarray, cells: np.ndarray, boundary: np.ndarray):
    """Promote P1 simplices to P3 (triangle10 / tetra20).

    Each unique edge (lo, hi) gets two nodes, at 1/3 and 2/3 from lo, numbered
    in order of first appearance; a cell lists them from its own local edge's
    first vertex (tri_basis / tet_basis node order, basis.py:639-859).  Each
    face gets one centroid node: per cell for triangles, per unique face for
    tetrahedra.  A new node is a boundary node iff all vertices of its edge /
    face are; triangle centroids are interior."""
    b = cells.shape[1]
    loc = np.asarray(TRI_EDGES if b == 3 else TET_EDGES, dtype=np.int64)
    n = points.shape[0]
    e0, e1 = cells[:, loc[:, 0]], cells[:, loc[:, 1]]
    lo, hi = np.minimum(e0, e1).ravel(), np.maximum(e0, e1).ravel()
    uniq, first_idx, inv = np.unique(lo * n + hi, return_index=True, return_inverse=True)
    order = np.argsort(first_idx, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    ulo, uhi = uniq[order] // n, uniq[order] % n
    n_e = ulo.size
    third = np.empty((2 * n_e, points.shape[1]))
    third[0::2] = (2.0 * points[ulo] + points[uhi]) / 3.0       # 1/3 from lo
    third[1::2] = (points[ulo] + 2.0 * points[uhi]) / 3.0       # 2/3 from lo
    near_lo = n + 2 * rank[inv]
    forward = (e0.ravel() == lo)                                  # local edge runs lo -> hi
    first = np.where(forward, near_lo, near_lo + 1).reshape(cells.shape[0], -1)
    second = np.where(forward, near_lo + 1, near_lo).reshape(cells.shape[0], -1)
    edge_nodes = np.stack([first, second], 2).reshape(cells.shape[0], -1)
    ebnd = np.repeat(boundary[ulo] & boundary[uhi], 2)
    base = n + 2 * n_e
    if b == 3:
        face_pts = points[cells].mean(1)
        face_nodes = (base + np.arange(cells.shape[0]))[:, None]
        fbnd = np.zeros(cells.shape[0], bool)
    else:
        floc = np.array([(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)], np.int64)
        fverts = np.sort(cells[:, floc], axis=2).reshape(-1, 3)
        key = (fverts[:, 0] * n + fverts[:, 1]) * n + fverts[:, 2]
        fu, ffirst, finv = np.unique(key, return_index=True, return_inverse=True)
        forder = np.argsort(ffirst, kind="stable")
        frank = np.empty_like(forder)
        frank[forder] = np.arange(forder.size)
        uf = fverts[ffirst[forder]]
        face_pts = points[uf].mean(1)
        face_nodes = (base + frank[finv]).reshape(cells.shape[0], 4)
        fbnd = boundary[uf].all(1)
    new_points = np.concatenate([points, third, face_pts], 0)
    new_boundary = np.concatenate([boundary, ebnd, fbnd])
    new_cells = np.concatenate([cells, edge_nodes, face_nodes], 1)
    return new_points, new_cells, new_boundary


def _tensor_grid_arrays(dim: int, n: int, order: int, jitter: float, seed: Optional[int]):
    """Structured n^dim grid of quadrilaterals / hexahedra on [0, 1]^dim.

    The element vertices are jittered (interior ones only); order-2 nodes sit
    on the (2n+1)^dim grid at the averages of the surrounding vertices, so
    edges stay straight.  Node ids are x-fastest on that grid; each cell lists
    its nodes in the reference order (vertices lexicographic, then edge, face
    and cell nodes)."""
    from .element import tensor_nodes
    h = 1.0 / n
    m = n + 1
    grids = np.meshgrid(*[np.arange(m)] * dim, indexing="ij")[::-1]        # x-fastest numbering
    idx = [g.ravel() for g in grids]
    coarse = np.stack([i * h for i in idx], 1).astype(np.float64)
    cb = np.zeros(coarse.shape[0], bool)
    for i in idx:
        cb |= (i == 0) | (i == n)
    coarse = _jitter(coarse, ~cb, jitter * h, seed).reshape([m] * dim + [dim])   # [z][y][x] / [y][x]
    p = order
    f = p * n + 1
    fgrids = np.meshgrid(*[np.arange(f)] * dim, indexing="ij")[::-1]
    fidx = [g.ravel() for g in fgrids]                                       # [x, y(, z)] fine indices
    pts = np.zeros((fidx[0].size, dim))
    corners = 0
    for bits in itertools.product((0, 1), repeat=dim):
        cidx = [(fi + (p - 1) * b) // p for fi, b in zip(fidx, bits)]       # floor / ceil of fi / p
        pts += coarse[tuple(cidx[::-1])]
        corners += 1
    pts /= corners
    boundary = np.zeros(pts.shape[0], bool)
    for fi in fidx:
        boundary |= (fi == 0) | (fi == f - 1)
    etype = {(2, 1): "quad", (2, 2): "quad9", (3, 1): "hexahedron", (3, 2): "hexahedron27"}[(dim, order)]
    off = np.rint(tensor_nodes(etype) * p).astype(np.int64)                 # [B, dim]
    cgrids = np.meshgrid(*[np.arange(n)] * dim, indexing="ij")[::-1]
    cbase = [g.ravel() * p for g in cgrids]
    stride = [f ** d for d in range(dim)]
    cells = np.zeros((cbase[0].size, off.shape[0]), np.int64)
    for d in range(dim):
        cells += (cbase[d][:, None] + off[None, :, d]) * stride[d]
    return pts, cells, boundary, etype


def unit_square_quad_arrays(n: int, jitter: float = 0.0, seed: Optional[int] = 0, order: int = 1):
    """numpy (points, cells, boundary) of an n x n quadrilateral grid (quad / quad9)."""
    pts, cells, bnd, _ = _tensor_grid_arrays(2, n, order, jitter, seed)
    return pts, cells, bnd


def unit_cube_hex_arrays(n: int, jitter: float = 0.0, seed: Optional[int] = 0, order: int = 1):
    """numpy (points, cells, boundary) of an n^3 hexahedral grid (hexahedron / hexahedron27)."""
    pts, cells, bnd, _ = _tensor_grid_arrays(3, n, order, jitter, seed)
    return pts, cells, bnd


def _to_mesh(pts, cells, boundary, etype, dtype, device) -> Mesh:
    return Mesh(torch.from_numpy(pts).to(dtype=dtype, device=device),
                {etype: torch.from_numpy(cells).to(device)},
                torch.from_numpy(boundary).to(device))


def unit_square_tri(n: int, jitter: float = 0.0, seed: Optional[int] = 0, order: int = 1,
                    dtype: torch.dtype = torch.float64, device="cpu") -> Mesh:
    pts, cells, bnd = unit_square_tri_arrays(n, jitter, seed)
    etype = "triangle"
    if order == 2:
        pts, cells, bnd = quadratic_arrays(pts, cells, bnd)
        etype = "triangle6"
    elif order == 3:
        pts, cells, bnd = cubic_arrays(pts, cells, bnd)
        etype = "triangle10"
    return _to_mesh(pts, cells, bnd, etype, dtype, device)


def unit_cube_tet(n: int, jitter: float = 0.0, seed: Optional[int] = 0, order: int = 1,
                  dtype: torch.dtype = torch.float64, device="cpu") -> Mesh:
    pts, cells, bnd = unit_cube_tet_arrays(n, jitter, seed)
    etype = "tetra"
    if order == 2:
        pts, cells, bnd = quadratic_arrays(pts, cells, bnd)
        etype = "tetra10"
    elif order == 3:
        pts, cells, bnd = cubic_arrays(pts, cells, bnd)
        etype = "tetra20"
    return _to_mesh(pts, cells, bnd, etype, dtype, device)


def unit_square_quad(n: int, jitter: float = 0.0, seed: Optional[int] = 0, order: int = 1,
                     dtype: torch.dtype = torch.float64, device="cpu") -> Mesh:
    pts, cells, bnd, etype = _tensor_grid_arrays(2, n, order, jitter, seed)
    return _to_mesh(pts, cells, bnd, etype, dtype, device)


def unit_cube_hex(n: int, jitter: float = 0.0, seed: Optional[int] = 0, order: int = 1,
                  dtype: torch.dtype = torch.float64, device="cpu") -> Mesh:
    pts, cells, bnd, etype = _tensor_grid_arrays(3, n, order, jitter, seed)
    return _to_mesh(pts, cells, bnd, etype, dtype, device)


def lognormal_coefficient(n: int, seed: int, sigma: float = 1.0) -> np.ndarray:
    """Per-element coefficient ``exp(z)``, ``z ~ N(0, sigma^2)``, seeded."""
    rng = np.random.default_rng(seed)
    return np.exp(sigma * rng.standard_normal(n))
