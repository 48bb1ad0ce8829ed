"""Recognition of built-in weak forms.

The reference hands any ``forward`` to ``torch.vmap``
(tensormesh/assemble/element_assembler.py:320-386).  Here a form is sent to
the HIP kernels when it *is* one of the built-in integrands, checked two
ways: its parameter names must match a known signature, and evaluating it on
random inputs must reproduce the built-in integrand to 1e-12.  Anything else
-- or anything that raises while probed -- keeps the generic torch path.

Recognised bilinear integrands (per quadrature point):

===================  ===========================  ===============================
form id              signature                    value
===================  ===========================  ===============================
laplace              (gradu, gradv)               gradu @ gradv        builtin.py:61-62
mass                 (u, v)                       u * v                builtin.py:99-100
diffusion_elem       (gradu, gradv, k), k in      k * (gradu @ gradv)  thermal_topology.py:48-52
                     element_data, shape [E]
diffusion_node       (gradu, gradv, k), k in      k * (gradu @ gradv)  forms.rst:172-184
                     point_data, shape [N]
linear_elasticity    (gradu, gradv)               Ba^T C(E, nu) Bb     builtin.py:172-178
===================  ===========================  ===============================

Recognised linear integrands: ``(v, f)`` -> ``f * v`` with ``f`` nodal
(README.md:103-118) and ``(v)`` -> ``c * v`` (builtin.py:730-770).
"""

from __future__ import annotations

import inspect
from dataclasses import dataclass
from typing import Callable, Optional

import torch


@dataclass
class FormMatch:
    form: str                    # laplace | mass | diffusion_elem | diffusion_node | linear_elasticity
    coeff_key: Optional[str] = None
    young: float = 1.0
    poisson: float = 0.3


def _close(a, b) -> bool:
    if not torch.is_tensor(a):
        a = torch.as_tensor(a, dtype=torch.float64)
    a = a.detach().to(torch.float64)
    if a.shape != b.shape:
        return False
    scale = max(float(b.abs().max()), 1e-300)
    return bool(((a - b).abs().max() <= 1e-12 * scale).item())


def _lame(E, nu):
    return E / (2.0 * (1.0 + nu)), E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))


def _elasticity_block(g1, g2, E, nu):
    mu, lam = _lame(E, nu)
    d = g1.shape[0]
    return lam * torch.outer(g1, g2) + mu * torch.outer(g2, g1) + mu * torch.dot(g1, g2) * torch.eye(d, dtype=g1.dtype)


def _call(fn: Callable, names, values: dict):
    return fn(*[values[n] for n in names])


def match_bilinear(fn: Callable, dim: int, element_data: dict, point_data: dict, scalar_data: dict,
                   owner=None) -> Optional[FormMatch]:
    try:
        names = list(inspect.signature(fn).parameters)
    except (TypeError, ValueError):
        return None
    gen = torch.Generator().manual_seed(1234)
    rnd = lambda *s: torch.rand(*s, generator=gen, dtype=torch.float64) * 2.0 - 1.0
    try:
        with torch.no_grad():
            if sorted(names) == ["gradu", "gradv"]:
                ok_lap, ok_el = True, True
                E = float(getattr(owner, "E", 1.0)) if owner is not None else 1.0
                nu = float(getattr(owner, "nu", 0.3)) if owner is not None else 0.3
                for _ in range(3):
                    g1, g2 = rnd(dim), rnd(dim)
                    out = _call(fn, names, {"gradu": g1, "gradv": g2})
                    if not torch.is_tensor(out):
                        out = torch.as_tensor(out)
                    if out.dim() == 0:
                        ok_el = False
                        ok_lap &= _close(out, torch.dot(g1, g2))
                    elif out.shape == (dim, dim):
                        ok_lap = False
                        ok_el &= _close(out, _elasticity_block(g1, g2, E, nu))
                    else:
                        return None
                if ok_lap:
                    return FormMatch("laplace")
                if ok_el:
                    return FormMatch("linear_elasticity", young=E, poisson=nu)
                return None
            if sorted(names) == ["u", "v"]:
                for _ in range(3):
                    a, b = rnd(()), rnd(())
                    if not _close(_call(fn, names, {"u": a, "v": b}), a * b):
                        return None
                return FormMatch("mass")
            if len(names) == 3 and "gradu" in names and "gradv" in names:
                key = [n for n in names if n not in ("gradu", "gradv")][0]
                if key in element_data:
                    kind = "diffusion_elem"
                    data = element_data[key]
                    if isinstance(data, dict):
                        if len(data) != 1:
                            return None
                        data = next(iter(data.values()))
                elif key in point_data and key not in ("x",):
                    kind = "diffusion_node"
                    data = point_data[key]
                else:
                    return None
                if not torch.is_tensor(data) or data.dim() != 1:
                    return None
                for _ in range(3):
                    g1, g2, k = rnd(dim), rnd(dim), rnd(())
                    if not _close(_call(fn, names, {"gradu": g1, "gradv": g2, key: k}), k * torch.dot(g1, g2)):
                        return None
                return FormMatch(kind, coeff_key=key)
    except Exception:
        return None
    return None


@dataclass
class LoadMatch:
    key: Optional[str]          # nodal field name, or None for a constant
    const: float = 1.0


def match_linear(fn: Callable, point_data: dict, scalar_data: dict) -> Optional[LoadMatch]:
    try:
        names = list(inspect.signature(fn).parameters)
    except (TypeError, ValueError):
        return None
    gen = torch.Generator().manual_seed(4321)
    rnd = lambda *s: torch.rand(*s, generator=gen, dtype=torch.float64) * 2.0 - 1.0
    try:
        with torch.no_grad():
            if names == ["v"]:
                a = rnd(())
                c = _call(fn, names, {"v": a})
                c = float(torch.as_tensor(c, dtype=torch.float64)) / float(a)
                for _ in range(3):
                    b = rnd(())
                    if not _close(_call(fn, names, {"v": b}), torch.tensor(c * float(b), dtype=torch.float64)):
                        return None
                return LoadMatch(None, c)
            if len(names) == 2 and "v" in names:
                key = [n for n in names if n != "v"][0]
                if key not in point_data or key == "x":
                    return None
                data = point_data[key]
                if not torch.is_tensor(data) or data.dim() != 1:
                    return None
                for _ in range(3):
                    v, f = rnd(()), rnd(())
                    if not _close(_call(fn, names, {"v": v, key: f}), f * v):
                        return None
                return LoadMatch(key)
    except Exception:
        return None
    return None
