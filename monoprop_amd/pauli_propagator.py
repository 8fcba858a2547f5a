"""Qubit front-end (mirror of reference ``src/monoprop/pauli_propagator.py``)."""

from __future__ import annotations

from collections.abc import Sequence

from .circuit import Circuit, ExpGate
from .conversion import slots_to_pauli
from .monomial_propagator import MonomialPropagator, ParameterValues
from .pauli import Pauli, PauliOperator


class PauliPropagator(MonomialPropagator[PauliOperator]):
    """Classical simulator for qubit (Pauli) observables; the cutoff is Pauli weight."""

    def __init__(
        self,
        initial_operator: PauliOperator,
        initial_state,
        *,
        cutoff: int,
        schrodinger_cutoff: int | None = None,
        lower_atol: float | None = None,
        upper_atol: float | None = None,
        comm=None,
    ) -> None:
        schrodinger_cutoff = None if schrodinger_cutoff is None else 2 * schrodinger_cutoff
        self._init_simulator(
            initial_operator.get_local_operator(),
            initial_state,
            cutoff=cutoff,
            schrodinger_cutoff=schrodinger_cutoff,
            cutoff_type="support",
            lower_atol=lower_atol,
            upper_atol=upper_atol,
            basis_change=None,
            comm=comm,
            basis="pauli",
        )

    @property
    def num_qubits(self) -> int:
        return self._system_size

    def evolved_operator(self, parameters: ParameterValues = None, *, atol: float = 1e-12) -> PauliOperator:
        raw = self._simulator.evolved_operator(self._bind(parameters), atol)
        terms = {Pauli(*slots_to_pauli(slots)): coeff for slots, coeff in raw.items()}
        return PauliOperator(terms, self.num_qubits)

    def _circuit_gates(self, circuit: Circuit) -> Sequence[ExpGate]:
        if circuit.family == "majorana":
            raise TypeError(
                "PauliPropagator requires a qubit circuit; its gates are Majorana/fermionic. "
                "Use MajoranaPropagator for those."
            )
        return circuit.gates

    def update_initial_operator(self, new_operator: PauliOperator) -> None:
        """Replace the initial operator's coefficients (reference MonomialPropagator.inl:353-377): in the
        Heisenberg picture every term must already be in the operator (RuntimeError otherwise) and the
        graph is kept; in the Schrodinger picture the observable is replaced.  Functionals created
        earlier raise from then on."""
        self._simulator.update_initial_operator(new_operator.get_local_operator().terms)
