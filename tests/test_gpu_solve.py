"""Native batched Jacobi-PCG against dense solves."""
import numpy as np
import pytest
import torch

import tensormesh_amd as tm
from tensormesh_amd.mesh import lognormal_coefficient, unit_square_tri
from tensormesh_amd.solve import BatchedPCG

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _system(n, seed):
    m = unit_square_tri(n, 0.2, seed).to(DEV)
    E = m.elements().shape[0]
    A = tm.DiffusionElementAssembler.from_mesh(m)
    K = A(element_data={"kappa": torch.from_numpy(lognormal_coefficient(E, seed)).to(DEV)})
    cond = tm.Condenser(m.boundary_mask)
    b = torch.ones(m.n_points, dtype=torch.float64, device=DEV)
    return cond(K, b)


@pytest.fixture(params=[1, 2, 5], ids=["one_launch", "two_launch", "five_launch"])
def variant(request, monkeypatch):
    monkeypatch.setattr(BatchedPCG, "launches", request.param)
    return request.param


@pytest.mark.parametrize("graph", [True, False])
def test_single_system_matches_dense(graph, variant):
    Ki, fi = _system(12, 3)
    s = BatchedPCG(Ki.row_ptr, Ki.col_indices, Ki.values, [0, Ki.shape[0]], rtol=1e-12)
    x, it, rel = s.solve(fi, graph=graph, chunk=16)
    ref = torch.linalg.solve(Ki.to_dense(), fi)
    assert rel[0].item() <= 1e-12
    assert torch.allclose(x, ref, rtol=1e-9, atol=1e-12)
    assert 0 < it[0].item() < 2000


def test_batched_uneven_systems_match_dense(variant):
    systems = [_system(n, n) for n in (5, 9, 14, 7)]
    rows, cols, vals, bs, offs = [], [], [], [], [0]
    rp = [0]
    for Ki, fi in systems:
        n = Ki.shape[0]
        rows.append(Ki.row_indices + offs[-1])
        cols.append(Ki.col_indices + offs[-1])
        vals.append(Ki.values)
        bs.append(fi)
        rp.append(rp[-1] + Ki.nnz)
        offs.append(offs[-1] + n)
    row = torch.cat(rows)
    row_ptr = torch.zeros(offs[-1] + 1, dtype=torch.int32, device=DEV)
    row_ptr[1:] = torch.cumsum(torch.bincount(row, minlength=offs[-1]), 0).to(torch.int32)
    s = BatchedPCG(row_ptr, torch.cat(cols), torch.cat(vals), offs, rtol=1e-11)
    x, it, rel = s.solve(torch.cat(bs), chunk=8)
    assert (rel <= 1e-11).all()
    for k, (Ki, fi) in enumerate(systems):
        ref = torch.linalg.solve(Ki.to_dense(), fi)
        assert torch.allclose(x[offs[k]:offs[k + 1]], ref, rtol=1e-8, atol=1e-12)
    # systems converge independently: different iteration counts
    assert len(set(it.tolist())) > 1


def test_sparse_matrix_solve_and_spmv_on_device():
    Ki, fi = _system(10, 4)
    u = Ki.solve(fi, rtol=1e-12)
    assert torch.allclose(Ki @ u, fi, rtol=1e-9, atol=1e-11)
    dense = Ki.to_dense()
    x = torch.randn(Ki.shape[0], dtype=torch.float64, device=DEV)
    assert torch.allclose(Ki @ x, dense @ x, rtol=1e-12, atol=1e-13)


def test_variants_agree_on_a_many_workgroup_system():
    """A 14k-row system spans 55 workgroups: the one- and two-launch
    variants re-sum their system's partials in every workgroup and produce
    bit-identical iterates; the five-launch variant is the same Krylov method
    in other recurrences, so iteration counts agree within a few and
    solutions to the tolerance."""
    Ki, fi = _system(120, 5)
    out = {}
    for launches in (1, 2, 5):
        BatchedPCG.launches = launches
        try:
            s = BatchedPCG(Ki.row_ptr, Ki.col_indices, Ki.values, [0, Ki.shape[0]], rtol=1e-10)
            out[launches] = s.solve(fi, chunk=32)
        finally:
            BatchedPCG.launches = 2
    x1, it1, r1 = out[1]
    assert torch.equal(out[1][0], out[2][0]) and torch.equal(out[1][1], out[2][1])
    (x2, it2, r2), (x5, it5, r5) = out[2], out[5]
    assert r2[0].item() <= 1e-10 and r5[0].item() <= 1e-10
    assert abs(it2[0].item() - it5[0].item()) <= max(3, it5[0].item() // 50)
    res = torch.linalg.norm(Ki @ x2 - fi) / torch.linalg.norm(fi)
    assert res.item() <= 2e-10
    assert torch.allclose(x2, x5, rtol=1e-7, atol=1e-12 * x5.abs().max().item())
    # deterministic: a second solve reproduces the iterate bit for bit
    BatchedPCG.launches = 2
    s = BatchedPCG(Ki.row_ptr, Ki.col_indices, Ki.values, [0, Ki.shape[0]], rtol=1e-10)
    x2b, _, _ = s.solve(fi, chunk=32)
    assert torch.equal(x2, x2b)


def test_two_launch_beyond_register_partials_matches_one_launch(monkeypatch):
    """A 531,441-row system spans 2,076 workgroups, more than the 2,048 whose
    partial sums the two-launch kernels load into registers ahead of use: the
    rest are added by the remainder loop, in the same order as the one-launch
    kernels' sums, so the iterates stay bit-identical and reach the tolerance
    (the library's own choice stops at 2,048 workgroups per system; a C-ABI
    caller may go further)."""
    Ki, fi = _system(730, 6)
    assert Ki.shape[0] > 2048 * 256
    monkeypatch.setattr(BatchedPCG, "TWO_LAUNCH_MAX_BLOCKS", 4096)
    out = {}
    for launches in (1, 2):
        monkeypatch.setattr(BatchedPCG, "launches", launches)
        s = BatchedPCG(Ki.row_ptr, Ki.col_indices, Ki.values, [0, Ki.shape[0]], rtol=1e-6)
        assert s.n_blocks > 2048 and s.variant == {1: 2, 2: 1}[launches]
        out[launches] = s.solve(fi, chunk=64)
    (x1, it1, r1), (x2, it2, r2) = out[1], out[2]
    assert torch.equal(x1, x2) and torch.equal(it1, it2)
    assert r2[0].item() <= 1e-6
    res = torch.linalg.norm(Ki @ x2 - fi) / torch.linalg.norm(fi)
    assert res.item() <= 2e-6
