"""monoprop_amd: Heisenberg operator propagation (Pauli and Majorana bases) on one AMD MI355X.

Drop-in for the propagate path of Algorithmiq's monoprop: ``PauliPropagator`` / ``MajoranaPropagator``
with ``from_circuit``, ``propagate``, ``expectation_value``, ``evolved_operator`` and the graph/gradient
API, backed by hand-written HIP kernels behind a C ABI (``include/monoprop_amd.h``).

The module attributes of the reference's compiled core (``src/monoprop/__init__.py:21-29``,
``bindings/bindings.cpp.in:111-130``) are provided with this build's values: ``MAX_NUM_MODES`` is the
reference's limit of 250 logical modes (the engine's rows are 256 modes wide), ``has_mpi`` is False (one GPU holds the whole operator), ``__variant__`` names
the HIP target.
"""

from ._engine import MAX_LOGICAL_NUM_MODES as _MAX_LOGICAL_NUM_MODES
from .exceptions import NumberOfModesInvalidError
from .circuit import Circuit, ExpGate, expand_monomials, validate_parameter_mapping
from .majorana import Majorana, MajoranaOperator
from .majorana_propagator import MajoranaPropagator
from .monomial_propagator import MonomialPropagator
from .pauli import Pauli, PauliOperator
from .pauli_propagator import PauliPropagator
from .utils import antihermitian_generator_correction, is_antihermitian, jordan_wigner_basis_change

__version__ = "0.1.0"
MAX_NUM_MODES = _MAX_LOGICAL_NUM_MODES
has_mpi = False
__variant__ = "hip-gfx950"
__build_type__ = "Release"


def __getattr__(name):
    # the build flags are read from the build script on first use, so that running it as
    # ``python -m monoprop_amd.build`` does not import it twice
    if name == "__compiler_flags__":
        from .build import FLAGS

        return " ".join(FLAGS)
    raise AttributeError(f"module 'monoprop_amd' has no attribute {name!r}")


__all__ = [
    "MAX_NUM_MODES",
    "Circuit",
    "ExpGate",
    "Majorana",
    "MajoranaOperator",
    "MajoranaPropagator",
    "MonomialPropagator",
    "NumberOfModesInvalidError",
    "Pauli",
    "PauliOperator",
    "PauliPropagator",
    "__build_type__",
    "__compiler_flags__",
    "__variant__",
    "__version__",
    "antihermitian_generator_correction",
    "expand_monomials",
    "has_mpi",
    "is_antihermitian",
    "jordan_wigner_basis_change",
    "validate_parameter_mapping",
]
