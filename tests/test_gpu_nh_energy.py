"""Neo-Hookean energy on the HIP kernel (tm_neohookean_energy) and its
backward (the HIP residual) against the oracle: energy within 1e-12
relative, energy(u).backward() against the oracle's assembled residual
(|d| <= 1e-11 |ref| + 1e-13 max|ref|), for simplex and tensor cells, both
with the residual formed in the forward launch and without."""
import sys

import pytest
import torch

import tensormesh_amd as tm
from oracle import fem
from tensormesh_amd.mesh import unit_cube_tet, unit_square_tri

pytestmark = pytest.mark.gpu
DEV = "cuda"


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _case(etype):
    order = 2 if etype in ("tetra10", "triangle6") else 1
    if etype.startswith("tri"):
        m = unit_square_tri(4, 0.1, 2, order=order)
    else:
        m = unit_cube_tet(3, 0.1, 2, order=order)
    P = m.points
    g = torch.Generator().manual_seed(9)
    cols = [torch.sin(P[:, 1]), P[:, 0] * P[:, -1], 0.5 * P[:, 0] ** 2][:P.shape[1]]
    u = 0.05 * torch.stack(cols, 1) + 0.002 * torch.randn(P.shape, generator=g, dtype=torch.float64)
    return m, u


@pytest.mark.parametrize("etype", ["triangle", "triangle6", "tetra", "tetra10"])
def test_neohookean_energy_and_backward_against_oracle(etype):
    say(f"case neo-hookean energy {etype}")
    m, u = _case(etype)
    e_ref = fem.neohookean_energy(m.points, m.elements(), etype, 2, u, E=1.0, nu=0.3)
    res_e, _ = fem.neohookean_element_tangent(m.points, m.elements(), etype, 2, u, E=1.0, nu=0.3)
    D = m.points.shape[1]
    res = torch.zeros(m.n_points, D, dtype=torch.float64).index_add(0, m.elements().reshape(-1),
                                                                     res_e.reshape(-1, D))
    model = tm.NeoHookeanModel.from_mesh(m.to(DEV), quadrature_order=2, E=1.0, nu=0.3)
    with torch.no_grad():
        e0 = model.energy(u.to(DEV))
    assert e0.grad_fn is None and e0.device.type == "cuda"
    ud = u.to(DEV).clone().requires_grad_(True)
    e = model.energy(ud)
    assert "NHEnergy" in type(e.grad_fn).__name__
    rel = abs(e.item() - e_ref.item()) / abs(e_ref.item())
    say(f"  energy {e.item():.16e} vs oracle {e_ref.item():.16e}: rel {rel:.2e}")
    assert rel <= 1e-12 and abs(e0.item() - e.item()) <= 1e-15 * abs(e.item())
    (2.5 * e).backward()
    d = (ud.grad.cpu() - 2.5 * res).abs()
    scale = res.abs().max()
    say(f"  backward vs oracle residual: max|d|/max|ref| {(d.max() / (2.5 * scale)).item():.2e}")
    assert bool((d <= 2.5 * (1e-11 * res.abs() + 1e-13 * scale)).all())
    # the residual entry point gives the same gradient
    R = model.residual(u.to(DEV))
    assert torch.allclose(R.cpu(), res.reshape(-1), rtol=1e-11, atol=1e-13 * scale.item())


def test_neohookean_energy_zero_at_rest_and_float32():
    say("case neo-hookean energy at rest and float32")
    m, u = _case("tetra10")
    md = m.to(DEV)
    model = tm.NeoHookeanModel.from_mesh(md, quadrature_order=2, E=1.0, nu=0.3)
    assert model.energy(torch.zeros_like(md.points)).item() == 0.0
    m32 = tm.Mesh(m.points.float(), {"tetra10": m.elements()}).to(DEV)
    e_ref = fem.neohookean_energy(m.points.float().double(), m.elements(), "tetra10", 2, u.float().double())
    e32 = tm.NeoHookeanModel.from_mesh(m32, quadrature_order=2, E=1.0, nu=0.3).energy(u.float().to(DEV))
    assert e32.dtype == torch.float32
    rel = abs(e32.item() - e_ref.item()) / abs(e_ref.item())
    say(f"  float32 energy rel err {rel:.2e}")
    assert rel <= 1e-6


def test_neohookean_energy_full_size_cfg5():
    say("case neo-hookean energy + backward, 32^3x6 tetra10 at u = 0.1x")
    m = unit_cube_tet(32, 0.0, 0, order=2)
    u = torch.zeros_like(m.points)
    u[:, 0] = 0.1 * m.points[:, 0]
    e_ref = fem.neohookean_energy(m.points, m.elements(), "tetra10", 2, u, E=1.0, nu=0.3)
    md = m.to(DEV)
    model = tm.NeoHookeanModel.from_mesh(md, quadrature_order=2, E=1.0, nu=0.3)
    ud = u.to(DEV).requires_grad_(True)
    e = model.energy(ud)
    e.backward()
    rel = abs(e.item() - e_ref.item()) / abs(e_ref.item())
    say(f"  energy rel err {rel:.2e}")
    assert rel <= 1e-12
    R = model.residual(u.to(DEV))
    assert torch.equal(ud.grad.reshape(-1), R)
