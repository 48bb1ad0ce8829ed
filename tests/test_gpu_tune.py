"""Autotuned launch choices (tensormesh_amd.tune) on the device: the tuned
path gives the same assembly as the fixed paths, the choice is cached, and
the adaptive PCG chunk leaves the solution unchanged bit for bit."""
import json
import sys

import pytest
import torch

import tensormesh_amd as tm
from tensormesh_amd import tune
from tensormesh_amd.mesh import lognormal_coefficient, unit_cube_tet, unit_square_tri

pytestmark = pytest.mark.gpu
DEV = "cuda"


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture
def fresh_tuner(tmp_path, monkeypatch):
    monkeypatch.setenv("TENSORMESH_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(tune, "_TUNER", None)
    yield tmp_path / "tune.json"
    monkeypatch.setattr(tune, "_TUNER", None)


def test_p1_path_auto_matches_fixed_paths_and_is_cached(fresh_tuner):
    say("case tune p1_path auto")
    m = unit_cube_tet(12, 0.15, 0).to(DEV)
    kap = torch.from_numpy(lognormal_coefficient(m.elements().shape[0], 3)).to(DEV)
    ed = {"kappa": kap}
    out = {}
    for path in ("rows", "two_stage", "two_stage_rows", "incidence", "auto"):
        A = tm.DiffusionElementAssembler.from_mesh(m, quadrature_order=1)
        A.p1_path = path
        out[path] = A(element_data=ed).values
    ref = out["two_stage"]
    for path, v in out.items():
        assert ((v - ref).abs() <= 1e-14 * ref.abs().max()).all(), path
    assert torch.equal(out["two_stage_rows"], ref) and torch.equal(out["incidence"], ref)
    data = json.loads(fresh_tuner.read_text())
    keys = [k for k in data if "|p1_path|" in k]
    assert len(keys) == 1 and data[keys[0]]["choice"] in ("rows", "two_stage", "two_stage_rows", "incidence")
    say(f"  chose {data[keys[0]]['choice']}: {data[keys[0]]['ms']}")
    assert torch.equal(out["auto"], out[data[keys[0]]["choice"]])


def test_nh_path_auto_matches_slots(fresh_tuner):
    say("case tune nh_path auto")
    m = unit_cube_tet(3, 0.1, 2, order=2).to(DEV)
    P = m.points
    u = 0.05 * torch.stack([torch.sin(P[:, 1]), P[:, 0] * P[:, 2], 0.5 * P[:, 0] ** 2], 1)
    res = {}
    for path in ("slots", "rows", "auto"):
        model = tm.NeoHookeanModel.from_mesh(m, quadrature_order=2, E=1.0, nu=0.3)
        model.nh_path = path
        res[path] = model.residual_and_tangent(u)
    (Rs, Ks), (Rr, Kr), (Ra, Ka) = res["slots"], res["rows"], res["auto"]
    assert torch.equal(Ks.values, Kr.values) and torch.equal(Rs, Rr)
    assert torch.equal(Ka.values, Ks.values) and torch.equal(Ra, Rs)


def test_adaptive_pcg_chunk_leaves_solution_unchanged():
    say("case tune pcg chunk")
    m = unit_square_tri(48, 0.2, 1).to(DEV)
    kap = torch.from_numpy(lognormal_coefficient(m.elements().shape[0], 4)).to(DEV)
    K = tm.DiffusionElementAssembler.from_mesh(m)(element_data={"kappa": kap})
    b = tm.SourceNodeAssembler.from_mesh(m)(point_data={"f": torch.ones(m.n_points, device=DEV,
                                                                         dtype=torch.float64)})
    Ki, fi = tm.Condenser(m.boundary_mask)(K, b)
    from tensormesh_amd.solve import BatchedPCG
    s = BatchedPCG(Ki.row_ptr, Ki.col_indices, Ki.values, [0, Ki.shape[0]], rtol=1e-10)
    x32, it32, _ = s.solve(fi, chunk=32)
    x_a, it_a, _ = s.solve(fi)                      # chunk from the previous iteration count
    assert s._last_iters == int(it_a.max()) and torch.equal(it32, it_a)
    assert torch.equal(x32, x_a)
    say(f"  iterations {int(it_a.max())}, adaptive chunk {tune.pcg_chunk(int(it32.max()))}")


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_elements_per_wavefront_choices_are_bit_identical(dtype, tmp_path, monkeypatch):
    """Every admitted elements-per-wavefront count of the element-matrix,
    Neo-Hookean and J2 wave kernels gives the same bits; "auto" times them
    once and keeps one; a count the kernel does not admit is rejected."""
    import importlib
    from tensormesh_amd import _native as nat
    from tensormesh_amd.mesh import unit_cube_hex, unit_cube_tet
    monkeypatch.setenv("TENSORMESH_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    tune = importlib.import_module("tensormesh_amd.tune")
    monkeypatch.setattr(tune, "_TUNER", None)
    m10 = unit_cube_tet(3, 0.1, 2, order=2)
    m10 = tm.Mesh(m10.points.to(dtype), m10.cells).to(DEV)
    m4 = unit_cube_tet(4, 0.1, 2)
    m4 = tm.Mesh(m4.points.to(dtype), m4.cells).to(DEV)
    mh = unit_cube_hex(3, 0.1, 2)
    mh = tm.Mesh(mh.points.to(dtype), mh.cells).to(DEV)
    P = m10.points
    u10 = 0.03 * torch.stack([torch.sin(P[:, 1]), P[:, 0] * P[:, 2], 0.5 * P[:, 0] ** 2], 1)
    P4 = m4.points
    u4 = torch.stack([3e-3 * P4[:, 0], -1e-3 * P4[:, 1], 5e-4 * P4[:, 2]], 1)
    cases = [
        ("elasticity tetra10", tm.LinearElasticityElementAssembler.from_mesh(m10, quadrature_order=2, E=1.0, nu=0.3),
         lambda A: A().values, (1, 2)),
        ("elasticity hexahedron", tm.LinearElasticityElementAssembler.from_mesh(mh, quadrature_order=2, E=1.0, nu=0.3),
         lambda A: A().values, (1, 2)),
        ("neo-hookean tetra10", tm.NeoHookeanModel.from_mesh(m10, quadrature_order=2, E=1.0, nu=0.3),
         lambda A: torch.cat([t.reshape(-1) for t in (A.residual_and_tangent(u10)[0],
                                                     A.residual_and_tangent(u10)[1].values)]), (1, 2, 4)),
        ("neo-hookean energy", tm.NeoHookeanModel.from_mesh(m10, quadrature_order=2, E=1.0, nu=0.3),
         lambda A: torch.cat([A._native_energy(u10, True)[0].reshape(1), A._native_energy(u10, True)[1]]),
         (1, 2, 4)),
        ("j2 tetra", tm.J2Plasticity.from_mesh(m4, quadrature_order=2, E=200.0, nu=0.3, sig0=0.25, H=1.0),
         lambda A: torch.cat([A.residual_and_tangent(u4)[0], A.residual_and_tangent(u4)[1].values]), (1, 4, 8)),
    ]
    for name, A, run, options in cases:
        outs = {}
        for e in options:
            A.epw = e
            outs[e] = run(A)
        for e in options[1:]:
            assert torch.equal(outs[options[0]], outs[e]), f"{name}: epw {e} differs from {options[0]}"
        A.epw = "auto"
        assert torch.equal(run(A), outs[options[0]]), name
        A.epw = 3
        with pytest.raises(nat.NativeError):
            run(A)
        A.epw = None
    data = json.loads((tmp_path / "tune.json").read_text())
    assert any("|epw|neohookean|" in k for k in data) and any("|epw|j2|" in k for k in data)
