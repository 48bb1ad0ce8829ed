"""The fused P1 kernel with the contribution order (each entry a contiguous
sum, tm_rows_dest_build) against the same kernel walking the incidence masks:
bit-identical values (same local rows, same summation order), for every P1
form on triangles and tetrahedra, including the full 1.57M-tet workload."""
import ctypes
import sys

import pytest
import torch

import tensormesh_amd as tm
from oracle import fem
from tensormesh_amd import _native as nat
from tensormesh_amd.mesh import Mesh, lognormal_coefficient, unit_cube_tet, unit_cube_tet_arrays, unit_square_tri

pytestmark = pytest.mark.gpu
DEV = "cuda"


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _both_orders(A, points, coeff_key, coeff, form_id):
    dev = points.device
    pat = A._pattern_on(dev)
    conn = A._conn32(dev)
    tab, tab_dev = A._tables(dev)
    from tensormesh_amd.assemble import _n_quad
    form = nat.FormDesc(nat.FORM[form_id], _n_quad(A.etype, A.quadrature_order), tab.ctypes.data,
                        tab_dev.data_ptr(), None if coeff is None else coeff.data_ptr(), 0.0, 0.0)
    mesh = A._mesh_view(points)
    out = []
    for order in (True, False):
        rv = pat.rows_view(conn, order=order)
        assert rv is not None and (rv.dest is not None) == order
        vals = torch.full((pat.nnz,), float("nan"), dtype=points.dtype, device=dev)
        nat.check(nat.lib().tm_assemble_p1_rows(ctypes.byref(mesh), ctypes.byref(form), ctypes.byref(rv),
                                                vals.data_ptr(), nat.stream_of(dev)), "tm_assemble_p1_rows")
        out.append(vals)
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("kind", ["tri", "tet"])
@pytest.mark.parametrize("form", ["laplace", "diffusion_elem", "diffusion_node", "mass"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_contribution_order_equals_mask_walk(kind, form, dtype):
    say(f"case rows order {kind} {form} {dtype}")
    m = unit_square_tri(9, 0.2, 4) if kind == "tri" else unit_cube_tet(5, 0.15, 4)
    md = Mesh(m.points.to(dtype), m.cells, m.boundary_mask).to(DEV)
    A = tm.DiffusionElementAssembler.from_mesh(md, quadrature_order=2)
    coeff = None
    if form == "diffusion_elem":
        coeff = torch.from_numpy(lognormal_coefficient(m.elements().shape[0], 3)).to(DEV, dtype)
    elif form == "diffusion_node":
        coeff = torch.from_numpy(lognormal_coefficient(m.n_points, 3)).to(DEV, dtype)
    a, b = _both_orders(A, md.points, None, coeff, form)
    assert torch.equal(a, b)
    if dtype == torch.float64:
        ref = fem.assemble_bilinear(m.points, m.elements(), m.default_element_type, 2, form,
                                    kappa=None if coeff is None else coeff.cpu())[0]
        err = ((a.cpu() - ref).abs().max() / ref.abs().max()).item()
        assert err < 1e-13, err


def test_contribution_order_full_size_tets():
    say("case rows order, 64^3x6 Kuhn tets, diffusion fp64")
    p, c, _ = unit_cube_tet_arrays(64, 0.15, 0)
    md = Mesh(torch.from_numpy(p), {"tetra": torch.from_numpy(c)}).to(DEV)
    kap = torch.from_numpy(lognormal_coefficient(c.shape[0], 1)).to(DEV)
    A = tm.DiffusionElementAssembler.from_mesh(md, quadrature_order=1)
    a, b = _both_orders(A, md.points, "kappa", kap, "diffusion_elem")
    assert torch.equal(a, b)
