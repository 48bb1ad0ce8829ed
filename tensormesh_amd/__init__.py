"""tensormesh_amd: MI355X-native tensorized finite-element assembly.

Drop-in for TensorMesh's assembly path (ElementAssembler / NodeAssembler /
Condenser / SparseMatrix) with the built-in forms running on hand-written
gfx950 HIP kernels behind the C-ABI of libtmhip.so.
"""

from .assemble import ElementAssembler, NodeAssembler
from .builtin import (DiffusionElementAssembler, J2Plasticity, LaplaceElementAssembler,
                      LinearElasticityElementAssembler, MassElementAssembler, NeoHookeanModel, SourceNodeAssembler, const_node_assembler,
                      func_node_assembler)
from .condense import Condenser
from .contact import ContactAssembler, PenaltyContact
from .facet import FacetAssembler
from .mesh import Mesh
from .sparse import SparseMatrix

__version__ = "0.1.0"

__all__ = [
    "ElementAssembler", "NodeAssembler", "FacetAssembler", "ContactAssembler", "PenaltyContact", "Condenser", "SparseMatrix", "Mesh",
    "LaplaceElementAssembler", "DiffusionElementAssembler", "MassElementAssembler",
    "LinearElasticityElementAssembler", "NeoHookeanModel", "J2Plasticity", "SourceNodeAssembler",
    "const_node_assembler", "func_node_assembler",
]
