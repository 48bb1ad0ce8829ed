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

Why agreement at a few random points is enough: the probes run under
``PolynomialGuard``, which rejects the integrand (generic torch path) as soon
as it applies anything but additions, products, constant scalings and
reshaping to its arguments or to values derived from them -- comparisons,
``torch.where``, ``abs``, transcendental functions, division by an argument,
reading a value into Python (``item``, ``bool``, ``if``), ``detach`` or
``.data``.  An integrand that passes the guard is a polynomial of low degree
in its arguments' entries, and so is its difference from the built-in
integrand; a nonzero polynomial of degree d vanishes at a point drawn
uniformly from a set of size m with probability at most d/m
(Schwartz-Zippel), i.e. never, in practice, at three independent float64
draws.  Integrands with data-dependent branches therefore never reach the
kernels, however they behave at the probe points.

Recognition results are cached per integrand together with a fingerprint of
everything it can read besides its arguments (``free_state``): the values of
its closure cells, default arguments and referenced globals, the owning
module's scalar attributes and tensor attributes (identity and version
counter), following called Python functions a few levels deep.  Changing any
of these re-probes; an integrand whose free state cannot be fingerprinted is
probed on every call.  The caches are bounded (least recently used entries
are dropped).
"""

from __future__ import annotations

import collections
import inspect
import types
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np
import torch
from torch.overrides import TorchFunctionMode


# ------------------------------------------------------ polynomial guard ---
# torch functions / tensor methods that keep a value polynomial in the form's
# arguments when applied to argument-derived tensors
_POLY_OPS = frozenset("""
matmul __matmul__ __rmatmul__ dot vdot inner outer ger mm mv bmm einsum tensordot kron
mul __mul__ __rmul__ multiply mul_ __imul__
add __add__ __radd__ add_ __iadd__ sub __sub__ __rsub__ subtract sub_ __isub__ rsub
neg __neg__ negative neg_ __pos__ positive
sum mean trace diagonal diag diag_embed triu tril cumsum
reshape view view_as reshape_as flatten unflatten squeeze unsqueeze transpose t permute movedim moveaxis
swapaxes swapdims expand expand_as broadcast_to repeat tile contiguous clone
__getitem__ __setitem__ select narrow index_select gather take unbind split chunk __iter__ __len__
stack cat concat concatenate hstack vstack dstack column_stack
zeros_like ones_like full_like empty_like new_zeros new_ones new_full new_empty new_tensor
to type double float type_as copy_ fill_ zero_ square dim size numel
""".split())
_DIV_OPS = frozenset(["div", "__truediv__", "true_divide", "divide", "div_", "__itruediv__"])
_POW_OPS = frozenset(["pow", "__pow__", "pow_", "__ipow__"])
# tensor properties that read shape or layout, not values outside autograd
_PROPERTIES = frozenset(["T", "mT", "shape", "dtype", "device", "ndim", "is_cuda", "layout", "requires_grad",
                         "grad_fn", "is_leaf", "names"])


def _tensors(args, kwargs):
    for a in list(args) + list(kwargs.values()):
        if torch.is_tensor(a):
            yield a
        elif isinstance(a, (list, tuple)):
            yield from (x for x in a if torch.is_tensor(x))


class PolynomialGuard(TorchFunctionMode):
    """Records whether any operation on an argument-derived tensor (one that
    requires grad: the probe arguments do) falls outside the polynomial
    whitelist.  ``ok`` is False after the first such operation."""

    def __init__(self):
        super().__init__()
        self.ok = True
        self.offender = None

    def _fail(self, name):
        if self.ok:
            self.ok, self.offender = False, name

    def __torch_function__(self, func, types_, args=(), kwargs=None):
        kwargs = kwargs or {}
        if any(t.requires_grad for t in _tensors(args, kwargs)):
            name = getattr(func, "__name__", repr(func))
            if not torch.is_grad_enabled():
                self._fail(f"{name} without autograd")
            elif name == "__get__":
                prop = getattr(getattr(func, "__self__", None), "__name__", "?")
                if prop not in _PROPERTIES:
                    self._fail(f".{prop}")
            elif name in _DIV_OPS:
                den = args[1] if len(args) > 1 else kwargs.get("other")
                if torch.is_tensor(den) and den.requires_grad:
                    self._fail(name)
            elif name in _POW_OPS:
                ex = args[1] if len(args) > 1 else kwargs.get("exponent")
                if not (isinstance(ex, int) and not isinstance(ex, bool) and 0 <= ex <= 8):
                    self._fail(name)
            elif name not in _POLY_OPS:
                self._fail(name)
        return func(*args, **kwargs)


def probe(fn: Callable, args: dict, guard: PolynomialGuard):
    """``fn(**args)`` with every tensor argument a fresh leaf requiring grad,
    evaluated under ``guard``; returns the detached float64 result."""
    leaves = {k: (v.detach().clone().requires_grad_(True) if torch.is_tensor(v) else v) for k, v in args.items()}
    with torch.enable_grad(), guard:
        out = fn(**leaves)
    if not torch.is_tensor(out):
        out = torch.as_tensor(out, dtype=torch.float64)
    return out.detach().to(torch.float64)


# ------------------------------------------------------- free-state keys ---
class _Unkeyable(Exception):
    pass


_MAX_DEPTH = 3


def _value_key(v, depth: int, seen: set):
    if v is None or isinstance(v, (bool, int, float, complex, str, bytes)):
        return (type(v).__name__, v)
    if isinstance(v, (torch.dtype, torch.device)):
        return ("t", str(v))
    if torch.is_tensor(v):
        return ("tensor", id(v), v._version, v.data_ptr() if v.numel() else 0, tuple(v.shape), str(v.dtype),
                str(v.device))
    if isinstance(v, (tuple, list, frozenset)) and len(v) <= 32:
        return (type(v).__name__,) + tuple(_value_key(x, depth, seen) for x in v)
    if isinstance(v, dict) and len(v) <= 32:
        return ("dict",) + tuple(sorted((str(k), _value_key(x, depth, seen)) for k, x in v.items()))
    if isinstance(v, np.ndarray) and v.size <= 1024:
        return ("ndarray", v.dtype.str, v.shape, v.tobytes())
    if isinstance(v, type) and not v.__module__.startswith(("torch", "builtins", "numpy")):
        # user classes: their plain class attributes
        return ("class", id(v), tuple(sorted((k, x) for k, x in vars(v).items()
                                             if isinstance(x, (bool, int, float, str)))))
    if isinstance(v, (types.ModuleType, type, types.BuiltinFunctionType, types.BuiltinMethodType)):
        return ("id", id(v))
    if isinstance(v, torch.nn.Module):
        return ("module", id(v), _owner_key(v, depth, seen))
    if callable(v) and hasattr(v, "__code__") or isinstance(v, types.MethodType):
        return ("fn", free_state(v, depth + 1, seen))
    if callable(v) and getattr(type(v), "__module__", "").startswith("torch"):
        return ("id", id(v))                  # torch's own callables (ops, dispatchers)
    raise _Unkeyable(type(v).__name__)


def _owner_key(obj, depth: int, seen: set):
    if id(obj) in seen:
        return ("seen", id(obj))
    seen.add(id(obj))
    out = []
    for k, v in sorted(obj.__dict__.items()):
        if k.startswith("_") and k not in ("_parameters", "_buffers", "_modules"):
            continue
        if k in ("_parameters", "_buffers"):
            out.extend((f"{k}.{n}", _value_key(t, depth, seen)) for n, t in v.items() if t is not None)
        elif k == "_modules":
            out.extend((f"{k}.{n}", _value_key(m, depth + 1, seen)) for n, m in v.items() if m is not None)
        elif isinstance(v, (bool, int, float, str)) or torch.is_tensor(v):
            out.append((k, _value_key(v, depth, seen)))
    return tuple(out)


def free_state(fn: Callable, depth: int = 0, seen: Optional[set] = None):
    """Hashable fingerprint of what ``fn`` can read besides its arguments, or
    raises ``_Unkeyable``.  Covers bound methods (the owner's scalar and
    tensor attributes), closures, defaults and referenced globals, and the
    Python functions those name, ``_MAX_DEPTH`` levels deep."""
    if depth > _MAX_DEPTH:
        raise _Unkeyable("call depth")
    seen = set() if seen is None else seen
    owner = None
    if isinstance(fn, types.MethodType):
        owner, fn = fn.__self__, fn.__func__
    code = getattr(fn, "__code__", None)
    if code is None:
        raise _Unkeyable(type(fn).__name__)
    if id(fn) in seen:
        return ("seen", id(fn))
    seen.add(id(fn))
    cells = tuple(_value_key(c.cell_contents, depth, seen) for c in (fn.__closure__ or ()))
    defaults = tuple(_value_key(d, depth, seen) for d in (fn.__defaults__ or ()))
    kwdefaults = tuple(sorted((k, _value_key(d, depth, seen)) for k, d in (fn.__kwdefaults__ or {}).items()))
    g = fn.__globals__
    names = set(code.co_names)
    for const in code.co_consts:            # names read by nested lambdas / comprehensions
        if isinstance(const, types.CodeType):
            names.update(const.co_names)
    globs = tuple(sorted((n, _value_key(g[n], depth, seen)) for n in names if n in g))
    own = None if owner is None else _owner_key(owner, depth, seen)
    return (id(fn), own, cells, defaults, kwdefaults, globs)


class MatchCache:
    """Bounded least-recently-used cache of recognition results keyed by the
    integrand, its free state and the call's data layout."""

    def __init__(self, size: int = 32):
        self.size = size
        self._d = collections.OrderedDict()

    def get(self, fn: Callable, layout, compute: Callable):
        try:
            key = (free_state(fn), layout)
            hash(key)
        except (_Unkeyable, TypeError, ValueError, RuntimeError):
            return compute()                 # not fingerprintable: probe on every call
        if key in self._d:
            self._d.move_to_end(key)
            return self._d[key]
        val = compute()
        self._d[key] = val
        if len(self._d) > self.size:
            self._d.popitem(last=False)
        return val


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
    guard = PolynomialGuard()
    try:
        if sorted(names) == ["gradu", "gradv"]:
            ok_lap, ok_el = True, True
            E = float(getattr(owner, "E", 1.0)) if owner is not None else 1.0
            nu = float(getattr(owner, "nu", 0.3)) if owner is not None else 0.3
            for _ in range(3):
                g1, g2 = rnd(dim), rnd(dim)
                out = probe(fn, {"gradu": g1, "gradv": g2}, guard)
                if out.dim() == 0:
                    ok_el = False
                    ok_lap &= _close(out, torch.dot(g1, g2))
                elif out.shape == (dim, dim):
                    ok_lap = False
                    ok_el &= _close(out, _elasticity_block(g1, g2, E, nu))
                else:
                    return None
            if not guard.ok:
                return None
            if ok_lap:
                return FormMatch("laplace")
            if ok_el:
                return FormMatch("linear_elasticity", young=E, poisson=nu)
            return None
        if sorted(names) == ["u", "v"]:
            for _ in range(3):
                a, b = rnd(()), rnd(())
                if not _close(probe(fn, {"u": a, "v": b}, guard), a * b):
                    return None
            return FormMatch("mass") if guard.ok else None
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
                if not _close(probe(fn, {"gradu": g1, "gradv": g2, key: k}, guard), k * torch.dot(g1, g2)):
                    return None
            return FormMatch(kind, coeff_key=key) if guard.ok else None
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
    guard = PolynomialGuard()
    try:
        if names == ["v"]:
            a = rnd(())
            c = float(probe(fn, {"v": a}, guard)) / float(a)
            for _ in range(3):
                b = rnd(())
                if not _close(probe(fn, {"v": b}, guard), torch.tensor(c * float(b), dtype=torch.float64)):
                    return None
            return LoadMatch(None, c) if guard.ok else None
        if len(names) == 2 and "v" in names:
            key = [n for n in names if n != "v"][0]
            if key not in point_data or key == "x":
                return None
            data = point_data[key]
            if not torch.is_tensor(data) or data.dim() != 1:
                return None
            for _ in range(3):
                v, f = rnd(()), rnd(())
                if not _close(probe(fn, {"v": v, key: f}, guard), f * v):
                    return None
            return LoadMatch(key) if guard.ok else None
    except Exception:
        return None
    return None
