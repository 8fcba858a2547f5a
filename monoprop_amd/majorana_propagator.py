"""Fermionic front-end (mirror of reference ``src/monoprop/majorana_propagator.py``)."""

from __future__ import annotations

from collections.abc import Sequence

from .circuit import Circuit, ExpGate
from .majorana import MajoranaOperator
from .monomial_propagator import MonomialPropagator, ParameterValues


class MajoranaPropagator(MonomialPropagator[MajoranaOperator]):
    """Classical simulator for Majorana observables; cutoff by ``"length"`` (default) or ``"support"``."""

    def __init__(
        self,
        initial_operator,
        initial_state,
        *,
        cutoff: int,
        schrodinger_cutoff: int | None = None,
        cutoff_type: str = "length",
        lower_atol: float | None = None,
        upper_atol: float | None = None,
        comm=None,
    ) -> None:
        op = initial_operator if isinstance(initial_operator, MajoranaOperator) else initial_operator.get_majorana_operator()
        self._init_simulator(
            op,
            initial_state,
            cutoff=cutoff,
            schrodinger_cutoff=schrodinger_cutoff,
            cutoff_type=cutoff_type,
            lower_atol=lower_atol,
            upper_atol=upper_atol,
            basis_change=None,
            comm=comm,
        )

    def _circuit_gates(self, circuit: Circuit) -> Sequence[ExpGate]:
        if circuit.family == "pauli":
            raise TypeError(
                "MajoranaPropagator cannot consume a qubit circuit; its gates are Pauli. "
                "Use PauliPropagator for qubit circuits."
            )
        return circuit.gates

    def evolved_operator(self, parameters: ParameterValues = None, *, atol: float = 1e-12) -> MajoranaOperator:
        terms = self._simulator.evolved_operator(self._bind(parameters), atol)
        return MajoranaOperator(terms, self.num_modes)

    def update_initial_operator(self, new_operator: MajoranaOperator) -> None:
        """Replace the initial operator's coefficients (reference MonomialPropagator.inl:353-377): in the
        Heisenberg picture every term must already be in the operator (RuntimeError otherwise) and the
        graph is kept; in the Schrodinger picture the observable is replaced.  Functionals created
        earlier raise from then on."""
        self._simulator.update_initial_operator(new_operator.terms)

    @MonomialPropagator.cutoff_type.setter
    def cutoff_type(self, new_cutoff_type: str) -> None:
        self._simulator.cutoff_type = new_cutoff_type
