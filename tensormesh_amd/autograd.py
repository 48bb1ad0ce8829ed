"""Hand-written adjoint of the P1 diffusion assembly (filled in below)."""

from __future__ import annotations


def adjoint_supported(assembler, match) -> bool:
    return False


def assemble_with_adjoint(assembler, match, points, coeff):
    raise NotImplementedError
