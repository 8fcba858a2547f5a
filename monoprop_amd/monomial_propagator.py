"""Shared front-end of the Pauli and Majorana propagators (mirror of reference
``src/monoprop/monomial_propagator.py``): same methods, argument meaning and errors, backed by the HIP
engine through :mod:`monoprop_amd._engine`.
"""

from __future__ import annotations

import logging
from abc import ABC, abstractmethod
from collections.abc import Callable, Sequence
from typing import Generic, TypeVar

import numpy as np

from ._engine import dispatch
from .circuit import Circuit, ExpGate, expand_monomials, validate_parameter_mapping
from .majorana import MajoranaOperator
from .pauli import PauliOperator
from .utils import validate_basis_change

logger = logging.getLogger(__name__)

T_op = TypeVar("T_op", MajoranaOperator, PauliOperator)
ParameterValues = Circuit | Sequence[float] | np.ndarray | None


class MonomialPropagator(ABC, Generic[T_op]):
    """Abstract base: subclasses resolve their operator to engine slot tuples and call ``_init_simulator``.

    Heisenberg evolution consumes each call's gates back to front, so splitting a circuit across calls
    is not equivalent to one call.
    """

    def _init_simulator(
        self,
        majorana_operator: MajoranaOperator,
        initial_state,
        *,
        cutoff: int,
        schrodinger_cutoff: int | None,
        cutoff_type: str,
        lower_atol: float | None,
        upper_atol: float | None,
        basis_change,
        comm,
        basis: str = "majorana",
    ) -> None:
        num_modes = majorana_operator.num_modes
        logger.debug("_init_simulator. num_modes=%d, cutoff=%d, schrodinger_cutoff=%s", num_modes, cutoff,
                     schrodinger_cutoff)
        validate_basis_change(basis_change, num_modes)
        self._comm = comm
        self._n_params = 0
        self._system_size = num_modes
        self._initial_state = list(initial_state)
        self._simulator = dispatch(num_modes)(
            initial_operator=majorana_operator.terms,
            cutoff=cutoff,
            initial_state=list(initial_state),
            schrodinger_cutoff=schrodinger_cutoff,
            lower_atol=lower_atol,
            upper_atol=upper_atol,
            cutoff_type=cutoff_type,
            basis_change=basis_change,
            comm=comm,
            basis=basis,
        )

    @classmethod
    def from_circuit(cls, circuit: Circuit, initial_operator, **config):
        """Construct and propagate ``circuit`` in one step (no graph is stored)."""
        prop = cls(initial_operator, list(circuit.initial_state), **config)
        prop.propagate(circuit)
        return prop

    @abstractmethod
    def _circuit_gates(self, circuit: Circuit) -> Sequence[ExpGate]:
        raise NotImplementedError

    def _check_initial_state(self, circuit: Circuit) -> None:
        if circuit._state_given and sorted(circuit.initial_state) != sorted(self._initial_state):
            raise ValueError(
                f"circuit.initial_state {list(circuit.initial_state)} does not match the propagator's initial "
                f"state {list(self._initial_state)}. Construct the propagator with this circuit's initial state "
                "(or via from_circuit)."
            )

    def _check_circuit_width(self, circuit: Circuit) -> None:
        if circuit.system_size != self._system_size:
            raise ValueError(
                f"Circuit system_size={circuit.system_size} does not match propagator width {self._system_size}."
            )

    def _validate_only_rotate_len_k(self, k: int | None) -> None:
        if k is not None and not 0 < k <= 2 * self._system_size:
            raise ValueError(f"only_rotate_len_k={k} is out of range; must be 0 < k <= 2*num_qubits ")

    def build_graph(self, circuit: Circuit, *, seed_parameters: ParameterValues = None,
                    only_rotate_len_k: int | None = None) -> None:
        """Append a circuit to the propagation graph (coefficient-seeded at the circuit's own
        parameters on the first call)."""
        self._check_initial_state(circuit)
        self._check_circuit_width(circuit)
        self._validate_only_rotate_len_k(only_rotate_len_k)
        if seed_parameters is not None:
            seed = seed_parameters
        elif self.graph_layers == 0:
            seed = circuit.parameters
        else:
            seed = None
        gates = self._circuit_gates(circuit)
        mapping = [self._n_params + m for m in circuit.resolved_mapping]
        majoranas, gen_coeffs, per_monomial, gate_indices = expand_monomials(gates, mapping, self._system_size)
        bound = self._bind(seed) or None
        self._simulator.build_graph(majoranas, per_monomial, gen_coeffs, gate_indices, bound, only_rotate_len_k)
        self._n_params += circuit.n_parameters

    def propagate(self, circuit: Circuit, *, only_rotate_len_k: int | None = None) -> None:
        """Evolve and contract immediately at the circuit's own parameters, without storing a graph."""
        self._validate_only_rotate_len_k(only_rotate_len_k)
        self._check_initial_state(circuit)
        self._check_circuit_width(circuit)
        gates = self._circuit_gates(circuit)
        majoranas, gen_coeffs, mapping, _ = expand_monomials(gates, circuit.resolved_mapping, self._system_size)
        self._simulator.propagate(majoranas, mapping, gen_coeffs, self._bind(circuit.parameters), only_rotate_len_k)

    @property
    def n_parameters(self) -> int:
        return self._n_params

    @property
    def n_gates(self) -> int:
        return self._simulator.n_gates()

    @property
    def parameter_mapping(self) -> list[int]:
        return list(self._simulator.parameter_mapping)

    @parameter_mapping.setter
    def parameter_mapping(self, mapping: Sequence[int]) -> None:
        resolved = [int(m) for m in mapping]
        n_layers, n_gates = self.graph_layers, self.n_gates
        if len(resolved) == n_layers:
            validate_parameter_mapping(resolved, n_layers, "graph layers")
        elif len(resolved) == n_gates:
            validate_parameter_mapping(resolved, n_gates, "gates")
        else:
            raise ValueError(
                f"parameter_mapping has {len(resolved)} entries; expected {n_layers} (per graph layer) or "
                f"{n_gates} (per gate)."
            )
        self._simulator.parameter_mapping = resolved
        self._n_params = max(resolved, default=-1) + 1

    def expectation_value(self, parameters: ParameterValues = None) -> float:
        """<ref| O(theta) |ref> (replays the stored graph at ``parameters`` when there is one)."""
        return self._simulator.expectation_value(self._bind(parameters))

    def expectation_value_and_gradient(self, parameters: ParameterValues = None) -> tuple[float, np.ndarray]:
        value, grad = self._simulator.expectation_value_and_gradient(self._bind(parameters))
        return value, np.asarray(grad, dtype=np.float64)

    def gradient(self, parameters: ParameterValues = None) -> np.ndarray:
        return self.expectation_value_and_gradient(parameters)[1]

    def expectation_value_functional(self, pare_threshold: float | None = None) -> Callable[..., float]:
        fn = self._simulator.expectation_value_functional(pare_threshold)
        return lambda parameters=None: fn(self._bind(parameters))

    def expectation_value_and_gradient_functional(self, pare_threshold: float | None = None) -> Callable[..., tuple]:
        fn = self._simulator.expectation_value_and_gradient_functional(pare_threshold)

        def call(parameters=None):
            value, grad = fn(self._bind(parameters))
            return value, np.asarray(grad, dtype=np.float64)

        return call

    def expval(self, parameters: ParameterValues = None) -> float:
        return self.expectation_value(parameters)

    def grad(self, parameters: ParameterValues = None) -> np.ndarray:
        return self.gradient(parameters)

    def expval_and_grad(self, parameters: ParameterValues = None) -> tuple[float, np.ndarray]:
        return self.expectation_value_and_gradient(parameters)

    def expval_functional(self, pare_threshold: float | None = None):
        return self.expectation_value_functional(pare_threshold)

    def expval_and_grad_functional(self, pare_threshold: float | None = None):
        return self.expectation_value_and_gradient_functional(pare_threshold)

    def contract_partially(self, parameters: ParameterValues = None, *, inplace: bool = True) -> np.ndarray:
        """Evolved coefficients at ``parameters`` (core term excluded); ``inplace`` folds the graph in."""
        coeffs = np.asarray(self._simulator.contract_partially(self._bind(parameters), inplace))
        if inplace:
            self._n_params = max(self._simulator.parameter_mapping, default=-1) + 1
        return coeffs

    @abstractmethod
    def evolved_operator(self, parameters: ParameterValues = None, *, atol: float = 1e-12) -> T_op:
        """The evolved operator with terms ``|c| >= atol``, coefficients rounded to 1e-12."""

    @abstractmethod
    def update_initial_operator(self, new_operator: T_op) -> None:
        """Re-weight the initial operator's existing terms."""

    def size(self) -> int:
        """Number of distinct monomial terms stored."""
        return self._simulator.size()

    def graph_size(self) -> tuple[int, int]:
        return self._simulator.graph_size()

    @property
    def num_modes(self) -> int:
        return self._system_size

    @property
    def graph_layers(self) -> int:
        return self._simulator.graph_layers()

    @property
    def schrodinger(self) -> bool:
        return self._simulator.schrodinger

    @property
    def cutoff(self) -> int:
        return self._simulator.cutoff

    @cutoff.setter
    def cutoff(self, new_cutoff: int) -> None:
        self._simulator.cutoff = new_cutoff

    @property
    def lower_atol(self) -> float | None:
        return self._simulator.lower_atol

    @lower_atol.setter
    def lower_atol(self, v: float | None) -> None:
        self._simulator.lower_atol = v

    @property
    def upper_atol(self) -> float | None:
        return self._simulator.upper_atol

    @upper_atol.setter
    def upper_atol(self, v: float | None) -> None:
        self._simulator.upper_atol = v

    @property
    def cutoff_type(self) -> str:
        return self._simulator.cutoff_type

    def _bind(self, parameters: ParameterValues) -> list[float]:
        if isinstance(parameters, Circuit):
            parameters = parameters.parameters
        if parameters is None:
            return []
        return [float(v) for v in parameters]
