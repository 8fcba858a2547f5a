"""Pauli terms and operators (mirror of reference ``src/monoprop/pauli.py``)."""

from __future__ import annotations

import itertools
from collections.abc import Mapping, Sequence

import numpy as np

from .conversion import pauli_string_to_majorana, pauli_to_slots
from .majorana import MajoranaOperator
from .utils import validate_system_size

_VALID = frozenset("IXYZ")


class Pauli:
    """Pauli letters on specific qubits, canonicalized: identities dropped, sorted by qubit."""

    __slots__ = ("qubits", "string")

    def __init__(self, string: str, qubits: int | Sequence[int] | None = None) -> None:
        if qubits is None:
            qubits = range(len(string))
        elif isinstance(qubits, int):
            qubits = (qubits,)
        qs = tuple(int(q) for q in qubits)
        bad = set(string) - _VALID
        if bad:
            raise ValueError(f"Invalid characters in Pauli string: {bad}. Only I, X, Y, Z are allowed.")
        if len(string) != len(qs):
            raise ValueError(f"Pauli string {string!r} and qubits {qs} must have the same length.")
        if len(set(qs)) != len(qs):
            raise ValueError(f"Duplicate qubit indices in Pauli term: {qs}.")
        if any(q < 0 for q in qs):
            raise ValueError(f"Pauli qubit indices must be non-negative; got {qs}.")
        pairs = sorted((q, p) for q, p in zip(qs, string) if p != "I")
        self.qubits = tuple(q for q, _ in pairs)
        self.string = "".join(p for _, p in pairs)

    def __eq__(self, other: object) -> bool:
        return isinstance(other, Pauli) and (self.string, self.qubits) == (other.string, other.qubits)

    def __hash__(self) -> int:
        return hash((self.string, self.qubits))

    def __repr__(self) -> str:
        return f"Pauli({self.string!r}, {self.qubits})"


class PauliOperator:
    """A real-weighted sum of Pauli terms on ``num_qubits`` qubits.

    Keys are :class:`Pauli` terms or full-width strings (``"ZZI"`` means qubits 0..2).
    """

    def __init__(self, terms: Mapping[Pauli | str, float], num_qubits: int) -> None:
        acc: dict[Pauli, float] = {}
        for key, coeff in terms.items():
            p = key if isinstance(key, Pauli) else Pauli(key)
            val = np.real_if_close(coeff)
            if np.iscomplexobj(val):
                raise ValueError("Operator has complex terms")
            acc[p] = acc.get(p, 0.0) + float(val)
        self.terms: dict[Pauli, float] = acc
        self.num_qubits = validate_system_size(num_qubits, argument_name="num_qubits")
        for p in self.terms:
            if p.qubits and p.qubits[-1] >= self.num_qubits:
                raise ValueError(f"Pauli term {p} acts on a qubit index >= num_qubits={self.num_qubits}.")

    @classmethod
    def _from_terms(cls, strings, coefficients, num_qubits) -> PauliOperator:
        acc: dict[Pauli, float] = {}
        for s, c in zip(strings, coefficients, strict=True):
            p = s if isinstance(s, Pauli) else Pauli(s)
            acc[p] = acc.get(p, 0.0) + c
        return cls(acc, num_qubits)

    def __len__(self) -> int:
        return len(self.terms)

    def __repr__(self) -> str:
        body = ""
        if len(self) <= 8:
            body = ": " + ", ".join(f"{c}*{p!r}" for p, c in self.terms.items())
        return f"PauliOperator({len(self)} terms, {self.num_qubits} qubits{body})"

    def __eq__(self, other: object) -> bool:
        if not isinstance(other, PauliOperator):
            return NotImplemented
        return self.num_qubits == other.num_qubits and self.terms == other.terms

    __hash__ = None  # type: ignore[assignment]

    def isclose(self, other: object, rtol: float = 1e-05, atol: float = 1e-8) -> bool:
        if not isinstance(other, PauliOperator):
            raise TypeError(f"Cannot compare PauliOperator with {type(other).__name__}.")
        if self.num_qubits != other.num_qubits:
            return False
        keys = set(self.terms) | set(other.terms)
        return all(
            np.isclose(self.terms.get(k, 0.0), other.terms.get(k, 0.0), rtol=rtol, atol=atol) for k in keys
        )

    def get_majorana_operator(self) -> MajoranaOperator:
        """Jordan-Wigner image as a :class:`MajoranaOperator` over ``num_qubits`` modes."""
        monos, coeffs = [], []
        for p, c in self.terms.items():
            full = ["I"] * self.num_qubits
            for q, letter in zip(p.qubits, p.string):
                full[q] = letter
            mono, phase = pauli_string_to_majorana("".join(full))
            monos.append(mono)
            coeffs.append(phase * c)
        return MajoranaOperator._from_terms(monos, coeffs, self.num_qubits)

    def get_local_operator(self) -> MajoranaOperator:
        """The native slot encoding, carried in a :class:`MajoranaOperator` container (not JW)."""
        if self.num_qubits is None:
            raise ValueError("get_local_operator() needs num_qubits.")
        monos = [pauli_to_slots(p.string, p.qubits) for p in self.terms]
        return MajoranaOperator._from_terms(monos, list(self.terms.values()), self.num_qubits)

    def all_pairwise_commute(self) -> bool:
        """Two Pauli terms commute iff they differ (both non-identity) on an even number of qubits."""
        dicts = [dict(zip(p.qubits, p.string)) for p in self.terms]
        for a, b in itertools.combinations(dicts, 2):
            if sum(a[q] != b[q] for q in a.keys() & b.keys()) % 2:
                return False
        return True
