"""Write the small golden fixtures under tests/golden/ from the CPU oracle.

Run from the repository root:  python tests/golden/make_golden.py
Fixtures (float64, .npz, no pickles):
  cfg1_tri20.npz   20x20 jittered triangle mesh (seed 0, +-0.2h): Laplace
                   stiffness (qorder 2), load with f = 2 pi^2 sin(pi x) sin(pi y)
                   at the nodes, Dirichlet condensation on the 80 boundary nodes
  tet3_diffusion.npz  3^3 Kuhn tets (+-0.15h, seed 1): element-kappa and
                   node-kappa diffusion, mass (qorder 1 and 2)
  tet10_elastic.npz   1^3 cube -> tetra10: linear elasticity E=1, nu=0.3 (qorder 2)
"""
import math
import os
import sys

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import fem  # noqa: E402
from tensormesh_amd.mesh import (lognormal_coefficient, quadratic_arrays, unit_cube_tet_arrays,  # noqa: E402
                                 unit_square_tri_arrays)

OUT = os.path.dirname(os.path.abspath(__file__))


def cfg1():
    p, c, bnd = unit_square_tri_arrays(20, 0.2, 0)
    P, C = torch.from_numpy(p), torch.from_numpy(c)
    v, r, cc, _ = fem.assemble_bilinear(P, C, "triangle", 2, "laplace")
    f = 2 * math.pi ** 2 * torch.sin(math.pi * P[:, 0]) * torch.sin(math.pi * P[:, 1])
    b = fem.assemble_linear(P, C, "triangle", 2, f)
    mask = torch.from_numpy(bnd)
    cd = fem.condense(v, r, cc, P.shape[0], mask, None, b)
    kv, kr, kc, _ = cd["K_ii"]
    np.savez(os.path.join(OUT, "cfg1_tri20.npz"), points=p, cells=c, boundary=bnd, row=r.numpy(), col=cc.numpy(),
             K=v.numpy(), f=f.numpy(), b=b.numpy(), Kii=kv.numpy(), Kii_row=kr.numpy(), Kii_col=kc.numpy(),
             f_i=cd["f_i"].numpy())


def tet3():
    p, c, bnd = unit_cube_tet_arrays(3, 0.15, 1)
    P, C = torch.from_numpy(p), torch.from_numpy(c)
    kap_e = torch.from_numpy(lognormal_coefficient(c.shape[0], 7))
    kap_n = torch.from_numpy(lognormal_coefficient(p.shape[0], 8))
    out = dict(points=p, cells=c, kappa_e=kap_e.numpy(), kappa_n=kap_n.numpy())
    for q in (1, 2):
        v, r, cc, _ = fem.assemble_bilinear(P, C, "tetra", q, "diffusion_elem", kappa=kap_e)
        out[f"diff_e_q{q}"] = v.numpy()
        out["row"], out["col"] = r.numpy(), cc.numpy()
        out[f"diff_n_q{q}"] = fem.assemble_bilinear(P, C, "tetra", q, "diffusion_node", kappa=kap_n)[0].numpy()
        out[f"mass_q{q}"] = fem.assemble_bilinear(P, C, "tetra", q, "mass")[0].numpy()
    np.savez(os.path.join(OUT, "tet3_diffusion.npz"), **out)


def tet10():
    p, c, bnd = quadratic_arrays(*unit_cube_tet_arrays(1, 0.0, 0))
    P, C = torch.from_numpy(p), torch.from_numpy(c)
    v, r, cc, _ = fem.assemble_bilinear(P, C, "tetra10", 2, "linear_elasticity", E=1.0, nu=0.3)
    np.savez(os.path.join(OUT, "tet10_elastic.npz"), points=p, cells=c, values=v.numpy(), row=r.numpy(),
             col=cc.numpy())


if __name__ == "__main__":
    cfg1()
    tet3()
    tet10()
    print("fixtures written to", OUT)
