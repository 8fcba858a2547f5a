"""monoprop_amd: Heisenberg operator propagation (Pauli and Majorana bases) on one AMD MI355X.

Drop-in for the propagate path of Algorithmiq's monoprop: ``PauliPropagator`` / ``MajoranaPropagator``
with ``from_circuit``, ``propagate``, ``expectation_value``, ``evolved_operator`` and the graph/gradient
API, backed by hand-written HIP kernels behind a C ABI (``include/monoprop_amd.h``).
"""

from .circuit import Circuit, ExpGate, expand_monomials
from .majorana import Majorana, MajoranaOperator
from .majorana_propagator import MajoranaPropagator
from .monomial_propagator import MonomialPropagator
from .pauli import Pauli, PauliOperator
from .pauli_propagator import PauliPropagator

__all__ = [
    "Circuit",
    "ExpGate",
    "Majorana",
    "MajoranaOperator",
    "MajoranaPropagator",
    "MonomialPropagator",
    "Pauli",
    "PauliOperator",
    "PauliPropagator",
    "expand_monomials",
]
