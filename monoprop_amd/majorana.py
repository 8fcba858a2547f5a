"""Majorana monomials and operators (mirror of reference ``src/monoprop/majorana.py``)."""

from __future__ import annotations

import itertools
from collections.abc import Mapping, Sequence

import numpy as np

from .conversion import cancel_pairs, permutation_sign
from .utils import validate_system_size


class Majorana:
    """One Majorana monomial ``m_{i1} ... m_{iw}`` with strictly increasing, non-negative indices.

    ``Majorana.from_unsorted`` canonicalizes an arbitrary index sequence and returns the reordering
    sign (repeated indices cancel in pairs).
    """

    __slots__ = ("indices",)

    def __init__(self, *indices: int) -> None:
        idx = tuple(int(i) for i in indices)
        if any(i < 0 for i in idx):
            raise ValueError(f"Majorana indices must be non-negative; got {idx}.")
        if any(a >= b for a, b in zip(idx, idx[1:])):
            raise ValueError(f"Majorana indices must be distinct and sorted; got {idx}.")
        self.indices = idx

    @classmethod
    def from_unsorted(cls, *indices: int) -> tuple[Majorana, float]:
        idx = tuple(int(i) for i in indices)
        if any(i < 0 for i in idx):
            raise ValueError(f"Majorana indices must be non-negative; got {idx}.")
        return cls(*cancel_pairs(sorted(idx))), float(permutation_sign(idx))

    def __eq__(self, other: object) -> bool:
        return isinstance(other, Majorana) and self.indices == other.indices

    def __hash__(self) -> int:
        return hash(self.indices)

    def __repr__(self) -> str:
        return f"Majorana({', '.join(map(str, self.indices))})"


class MajoranaOperator:
    """A weighted sum of Majorana monomials over ``num_modes`` fermionic modes (``2*num_modes`` indices).

    Keys are :class:`Majorana` terms or index tuples; each is canonicalized (sorted with its
    anticommutation sign, repeated pairs cancelled) and duplicates are summed.
    """

    def __init__(self, terms: Mapping[Majorana | Sequence[int], complex], num_modes: int) -> None:
        keys = [k.indices if isinstance(k, Majorana) else tuple(k) for k in terms]
        self.num_modes = validate_system_size(num_modes, argument_name="num_modes")
        for k in keys:
            for i in k:
                if int(i) < 0:
                    raise ValueError(f"Majorana indices must be non-negative; got {tuple(k)}.")
            if k and max(int(i) for i in k) >= 2 * self.num_modes:
                raise ValueError(f"Majorana term {tuple(k)} acts on an index >= num_modes={self.num_modes}.")
        self.terms = self._accumulate(keys, list(terms.values()))

    @classmethod
    def _from_terms(
        cls, majoranas: Sequence[Sequence[int]], coefficients: Sequence[complex], num_modes: int
    ) -> MajoranaOperator:
        obj = cls.__new__(cls)
        obj.num_modes = validate_system_size(num_modes, argument_name="num_modes")
        obj.terms = cls._accumulate(majoranas, coefficients)
        return obj

    @staticmethod
    def _accumulate(majoranas, coefficients) -> dict[tuple[int, ...], complex]:
        acc: dict[tuple[int, ...], complex] = {}
        for mono, coeff in zip(majoranas, coefficients, strict=True):
            term, sign = Majorana.from_unsorted(*mono)
            acc[term.indices] = acc.get(term.indices, 0j) + complex(coeff) * sign
        return acc

    def get_majorana_operator(self) -> MajoranaOperator:
        return self

    def __len__(self) -> int:
        return len(self.terms)

    def __repr__(self) -> str:
        body = ""
        if len(self) <= 8:
            body = ": " + ", ".join(f"{c}*{list(k)}" for k, c in self.terms.items())
        return f"MajoranaOperator({len(self)} terms, {self.num_modes} modes{body})"

    def __eq__(self, other: object) -> bool:
        if not isinstance(other, MajoranaOperator):
            return NotImplemented
        return self.num_modes == other.num_modes and self.terms == other.terms

    __hash__ = None  # type: ignore[assignment]

    def isclose(self, other: object, rtol: float = 1e-05, atol: float = 1e-8) -> bool:
        if not isinstance(other, MajoranaOperator):
            raise TypeError(f"Cannot compare MajoranaOperator with {type(other).__name__}.")
        if self.num_modes != other.num_modes:
            return False
        keys = set(self.terms) | set(other.terms)
        return all(
            np.isclose(self.terms.get(k, 0), other.terms.get(k, 0), rtol=rtol, atol=atol) for k in keys
        )

    def all_pairwise_commute(self) -> bool:
        """Monomials with supports S1, S2 commute iff |S1||S2| - |S1 & S2| is even."""
        sets = [set(k) for k in self.terms]
        for a, b in itertools.combinations(sets, 2):
            if (len(a) * len(b) - len(a & b)) % 2:
                return False
        return True
