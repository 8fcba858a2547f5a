r"""Circuit authoring: :class:`ExpGate` = ``exp(+i theta H)`` and :class:`Circuit`.

Mirror of reference ``src/monoprop/circuit.py``: same classes, argument meaning and errors.  A gate
expands to one layer per generator term, ``(slots, gen_coeff)``; the engine rotates by
``alpha = theta * gen_coeff`` with ``cos(2 alpha)``/``sin(2 alpha)``.
"""

from __future__ import annotations

from collections.abc import Iterator, Sequence
from typing import Literal

from .conversion import pauli_to_slots
from .majorana import MajoranaOperator
from .pauli import PauliOperator
from .utils import validate_system_size

GateFamily = Literal["pauli", "majorana"]
CircuitFamily = Literal["pauli", "majorana", "empty"]

#: An imaginary residue above this marks a non-Hermitian generator (reference circuit.py:503-504).
GENERATOR_HERMITICITY_ATOL = 1e-9


def _is_identity_gate(gate: ExpGate) -> bool:
    return all(c == 0 for c in gate.generator.terms.values())


class ExpGate:
    """``exp(+i theta H)`` for a Hermitian generator operator of commuting terms."""

    __slots__ = ("_atol", "_structural", "family", "generator", "index")

    def __init__(self, generator, index: int | None = None, *, atol: float = 1e-8, _structural: bool = False):
        if isinstance(generator, PauliOperator):
            family: GateFamily = "pauli"
        elif isinstance(generator, MajoranaOperator):
            family = "majorana"
        elif hasattr(generator, "get_majorana_operator"):
            generator = generator.get_majorana_operator()
            family = "majorana"
        else:
            raise TypeError(
                "ExpGate generator must be a MajoranaOperator, PauliOperator, or FermiOperator "
                f"(an operator object carrying its mode/qubit count), not a bare term; got {type(generator).__name__}."
            )
        if isinstance(generator, PauliOperator):
            kept = PauliOperator({p: c for p, c in generator.terms.items() if abs(c) > atol}, generator.num_qubits)
        else:
            kept = MajoranaOperator({m: c for m, c in generator.terms.items() if abs(c) > atol}, generator.num_modes)
        if not kept.all_pairwise_commute():
            raise ValueError("The provided generator must be composed of commuting terms.")
        self.generator = kept
        self.index = None if index is None else int(index)
        self.family = family
        self._structural = _structural
        self._atol = atol

    @classmethod
    def _structural_gate(cls, generator: MajoranaOperator, index: int | None) -> ExpGate:
        return cls(generator, index=index, _structural=True)

    @classmethod
    def _with_index(cls, gate: ExpGate, index: int | None) -> ExpGate:
        return cls(gate.generator, index=index, atol=gate._atol, _structural=gate._structural)

    def __eq__(self, other: object) -> bool:
        if not isinstance(other, ExpGate):
            return NotImplemented
        return (self.generator, self.index, self.family, self._structural) == (
            other.generator,
            other.index,
            other.family,
            other._structural,
        )

    __hash__ = None  # type: ignore[assignment]

    def __repr__(self) -> str:
        return f"ExpGate({self.generator}, index={self.index})"

    @property
    def system_size(self) -> int:
        g = self.generator
        return g.num_qubits if isinstance(g, PauliOperator) else g.num_modes


def validate_parameter_mapping(mapping: Sequence[int], expected_len: int, unit: str = "gates") -> None:
    """Mapping must have ``expected_len`` entries covering ``0..max`` without gaps."""
    if len(mapping) != expected_len:
        raise ValueError(f"parameter_mapping has {len(mapping)} entries but there are {expected_len} {unit}.")
    if mapping and set(mapping) != set(range(max(mapping) + 1)):
        raise ValueError(
            f"parameter_mapping indices must be contiguous 0..n-1 with no gaps; got {sorted(set(mapping))}."
        )


class Circuit:
    """An ordered sequence of :class:`ExpGate` of one family, angle values and a reference state."""

    def __init__(
        self,
        gates: Sequence[ExpGate],
        system_size: int,
        parameters: Sequence[float] = (),
        initial_state: Sequence[int] | None = None,
    ) -> None:
        gates = tuple(gates)
        parameters = tuple(float(v) for v in parameters)
        self._state_given = initial_state is not None
        state = tuple(int(i) for i in initial_state or ())
        system_size = validate_system_size(system_size, argument_name="system_size")
        if len(set(state)) != len(state):
            raise ValueError("Duplicate indices in initial state")
        if any(i < 0 or i >= system_size for i in state):
            raise ValueError(f"initial_state entries must be in 0..{system_size - 1}; got {list(state)}.")
        for g in gates:
            if not isinstance(g, ExpGate):
                raise TypeError(f"Circuit gates must be ExpGate; got {type(g).__name__}.")
            if g.system_size != system_size:
                raise ValueError(
                    f"Gate generator width {g.system_size} does not match circuit system_size={system_size}."
                )
        if all(g.index is None for g in gates) and any(_is_identity_gate(g) for g in gates):
            keep = [i for i, g in enumerate(gates) if not _is_identity_gate(g)]
            if parameters and len(parameters) != len(gates):
                raise ValueError(f"parameters has {len(parameters)} values but the circuit has {len(gates)} gates.")
            if parameters:
                parameters = tuple(parameters[i] for i in keep)
            gates = tuple(gates[i] for i in keep)
        self.gates = gates
        self.parameters = parameters
        self.initial_state = state
        self.system_size = system_size
        fams = {g.family for g in gates}
        if len(fams) > 1:
            raise TypeError(
                "A circuit cannot mix qubit and Majorana/fermionic gates; build separate circuits per family."
            )
        self.family: CircuitFamily = fams.pop() if fams else "empty"
        _ = self.resolved_mapping
        if self.parameters and len(self.parameters) != self.n_parameters:
            raise ValueError(
                f"parameters has {len(self.parameters)} values but the circuit has {self.n_parameters} parameters."
            )

    def __eq__(self, other: object) -> bool:
        if not isinstance(other, Circuit):
            return NotImplemented
        return (self.gates, self.parameters, self.initial_state, self.system_size) == (
            other.gates,
            other.parameters,
            other.initial_state,
            other.system_size,
        )

    __hash__ = None  # type: ignore[assignment]

    def __repr__(self) -> str:
        return (
            f"Circuit(gates={self.gates!r}, parameters={self.parameters!r}, "
            f"initial_state={self.initial_state!r}, system_size={self.system_size!r})"
        )

    @property
    def resolved_mapping(self) -> tuple[int, ...]:
        idx = [g.index for g in self.gates]
        if all(i is None for i in idx):
            return tuple(range(len(self.gates)))
        if any(i is None for i in idx):
            raise ValueError(
                "Either every gate must set `index`, or none of them must (mixing an explicit index with "
                "the default would be ambiguous)."
            )
        mapping = tuple(int(i) for i in idx)  # type: ignore[arg-type]
        validate_parameter_mapping(mapping, len(self.gates), "gates")
        return mapping

    @property
    def n_parameters(self) -> int:
        m = self.resolved_mapping
        return max(m) + 1 if m else 0

    def __len__(self) -> int:
        return len(self.gates)

    def __iter__(self) -> Iterator[ExpGate]:
        return iter(self.gates)

    def __add__(self, other: Circuit) -> Circuit:
        if not isinstance(other, Circuit):
            return NotImplemented
        if "empty" not in (self.family, other.family) and self.family != other.family:
            raise TypeError(
                f"Cannot concatenate a {self.family}-family circuit with a {other.family}-family one; "
                "the gate families differ."
            )
        if self._state_given and other._state_given and self.initial_state != other.initial_state:
            raise ValueError("Cannot concatenate circuits with different initial states.")
        if self.system_size != other.system_size:
            raise ValueError(
                f"Cannot concatenate circuits with different system_size: {self.system_size} != {other.system_size}."
            )
        off = self.n_parameters
        left = tuple(ExpGate._with_index(g, i) for g, i in zip(self.gates, self.resolved_mapping))
        right = tuple(ExpGate._with_index(g, i + off) for g, i in zip(other.gates, other.resolved_mapping))
        state = self.initial_state if self._state_given else (other.initial_state if other._state_given else None)
        return Circuit(left + right, self.system_size, tuple(self.parameters) + tuple(other.parameters), state)

    @classmethod
    def from_dense_arrays(
        cls,
        majoranas: Sequence[Sequence[int]],
        gen_coeffs: Sequence[float],
        param_inds: Sequence[int],
        system_size: int,
        parameters: Sequence[float] = (),
        initial_state: Sequence[int] | None = None,
    ) -> Circuit:
        """Majorana circuit from per-monomial arrays; consecutive equal ``param_inds`` form one gate
        whose coefficients are already the structural ``g`` (reference circuit.py:406-477)."""
        system_size = validate_system_size(system_size, argument_name="system_size")
        gates: list[ExpGate] = []
        run_monos: list[tuple[int, ...]] = []
        run_coeffs: list[complex] = []
        run_index: int | None = None

        def flush() -> None:
            op = MajoranaOperator._from_terms(run_monos, run_coeffs, num_modes=system_size)
            gates.append(ExpGate._structural_gate(op, index=run_index))

        for mono, coeff, pidx in zip(majoranas, gen_coeffs, [int(p) for p in param_inds], strict=True):
            if run_monos and pidx != run_index:
                flush()
                run_monos, run_coeffs = [], []
            run_index = pidx
            m = tuple(int(i) for i in mono)
            if any(i < 0 for i in m):
                raise ValueError(f"Majorana indices must be non-negative; got {m}.")
            if m and max(m) >= 2 * system_size:
                raise ValueError(f"Majorana term {m} acts on an index >= 2*system_size={2 * system_size}.")
            run_monos.append(m)
            run_coeffs.append(complex(float(coeff)))
        if run_monos:
            flush()
        return cls(
            gates=tuple(gates),
            parameters=tuple(float(p) for p in parameters),
            initial_state=tuple(int(i) for i in initial_state or ()),
            system_size=system_size,
        )


def _real_generator_coefficient(mono: Sequence[int], value: complex) -> float:
    value = complex(value)
    if abs(value.imag) > GENERATOR_HERMITICITY_ATOL:
        raise ValueError(
            f"Gate generator is not Hermitian: monomial {tuple(mono)} yields the structural generator "
            f"coefficient {value}, whose imaginary part is not negligible."
        )
    return float(value.real)


def _antihermitian_gen_coeff(mono: Sequence[int], coeff: complex) -> float:
    """Structural ``g = Re(-c / i^{w(w-1)/2})`` of a Hermitian Majorana generator term."""
    w = len(mono)
    return _real_generator_coefficient(mono, -complex(coeff) / (1j) ** (w * (w - 1) / 2))


def gate_layers(gate: ExpGate, num_qubits: int) -> list[tuple[tuple[int, ...], float]]:
    """Expand one gate into ``(slots, gen_coeff)`` layers in application order."""
    gen = gate.generator
    if isinstance(gen, PauliOperator):
        out = []
        for p, c in gen.terms.items():
            if p.qubits and p.qubits[-1] >= num_qubits:
                raise ValueError(
                    f"Gate generator term {p} acts on a qubit index >= the system's num_qubits={num_qubits}."
                )
            slots = pauli_to_slots(p.string, p.qubits)
            out.append((slots, _real_generator_coefficient(slots, c)))
        return out
    if gate._structural:
        return [(m, _real_generator_coefficient(m, c)) for m, c in gen.terms.items()]
    return [(m, _antihermitian_gen_coeff(m, c)) for m, c in gen.terms.items()]


def expand_monomials(gates: Sequence[ExpGate], mapping: Sequence[int], num_qubits: int):
    """Per-layer ``(majoranas, gen_coeffs, parameter_mapping, gate_indices)`` for the engine.

    A gate whose terms all fell below its atol expands to the identity layer ``((), 0.0)``.
    """
    num_qubits = validate_system_size(num_qubits, argument_name="num_qubits")
    monos: list[tuple[int, ...]] = []
    coeffs: list[float] = []
    per_layer: list[int] = []
    gate_idx: list[int] = []
    for gi, (gate, p) in enumerate(zip(gates, mapping, strict=True)):
        for mono, g in gate_layers(gate, num_qubits) or [((), 0.0)]:
            monos.append(tuple(mono))
            coeffs.append(g)
            per_layer.append(p)
            gate_idx.append(gi)
    return monos, coeffs, per_layer, gate_idx
