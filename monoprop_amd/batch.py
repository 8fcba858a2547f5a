"""Batched propagation: many (observable, parameter set) pairs of one circuit in one device store.

The reference answers such a sweep with one ``MonomialPropagator`` per pair: construct it with the
observable (``MonomialPropagator.inl:73-188``), ``propagate`` the circuit at the pair's parameters
(``inl:692-733``) and read ``expectation_value`` (``MPFunctions.cpp:135-148,196-205``).  Here the pairs
are members of one store: every gate is one pair of kernel launches for all members, each term
carries its member and takes that member's rotation constants, and the row index keys on
(monomial, member).  A member's rows, in store order, are exactly the rows its own propagator would
hold, with the same coefficients bit for bit.

``BatchPropagator`` is one such store.  ``expectation_grid`` runs a whole (parameter sets x
observables) sweep through as many stores as device memory needs, halving a batch that outgrows it.
"""

from __future__ import annotations

from collections.abc import Sequence

import numpy as np

from ._engine import BatchSimulator
from .circuit import Circuit, expand_monomials
from .majorana import MajoranaOperator
from .pauli import PauliOperator

Operator = PauliOperator | MajoranaOperator


def _local_terms(op: Operator) -> tuple[str, dict, int]:
    """(basis, slot-tuple terms, modes) of an observable, as the single propagators encode it."""
    if isinstance(op, PauliOperator):
        local = op.get_local_operator()
        return "pauli", local.terms, local.num_modes
    if isinstance(op, MajoranaOperator):
        return "majorana", op.terms, op.num_modes
    raise TypeError(f"observables must be PauliOperator or MajoranaOperator, got {type(op).__name__}")


class BatchPropagator:
    """Members ``q = 0 .. len(operators) - 1``: observable ``operators[q]`` propagated through one
    circuit at parameter row ``q``.  Truncation keywords are those of ``PauliPropagator`` (support
    cutoff) or ``MajoranaPropagator`` (``cutoff_type``, default ``"length"``)."""

    def __init__(self, operators: Sequence[Operator], initial_state: Sequence[int] = (), *, cutoff: int,
                 lower_atol: float | None = None, upper_atol: float | None = None,
                 cutoff_type: str | None = None) -> None:
        if not len(operators):
            raise ValueError("a batch needs at least one observable")
        encoded = [_local_terms(op) for op in operators]
        bases = {b for b, _, _ in encoded}
        widths = {n for _, _, n in encoded}
        if len(bases) != 1 or len(widths) != 1:
            raise ValueError("the observables of one batch must share their basis and width")
        basis = bases.pop()
        if basis == "pauli":
            if cutoff_type not in (None, "support"):
                raise ValueError("Pauli observables take the support cutoff")
            cutoff_type = "support"
        elif cutoff_type is None:
            cutoff_type = "length"
        self._basis = basis
        self._num_modes = widths.pop()
        self._initial_state = [int(i) for i in initial_state]
        self._sim = BatchSimulator([t for _, t, _ in encoded], cutoff, self._initial_state, lower_atol=lower_atol,
                                   upper_atol=upper_atol, cutoff_type=cutoff_type,
                                   logical_num_modes=self._num_modes, basis=basis)

    @property
    def members(self) -> int:
        return self._sim.members

    @property
    def num_modes(self) -> int:
        return self._num_modes

    def propagate(self, circuit: Circuit, parameter_sets) -> None:
        """Evolve every member through ``circuit`` (Heisenberg, back to front) at its own row of
        ``parameter_sets`` (shape ``(members, circuit.n_parameters)``)."""
        if circuit.system_size != self._num_modes:
            raise ValueError(f"Circuit system_size={circuit.system_size} does not match the batch width "
                             f"{self._num_modes}.")
        fam = circuit.family
        if (self._basis == "pauli" and fam == "majorana") or (self._basis == "majorana" and fam == "pauli"):
            raise TypeError(f"a {self._basis} batch cannot consume a {fam} circuit")
        if circuit._state_given and sorted(circuit.initial_state) != sorted(self._initial_state):
            raise ValueError("circuit.initial_state does not match the batch's initial state")
        majoranas, gen_coeffs, mapping, _ = expand_monomials(circuit.gates, circuit.resolved_mapping,
                                                             self._num_modes)
        params = np.asarray(parameter_sets, dtype=np.float64)
        if params.ndim != 2 or params.shape != (self.members, circuit.n_parameters):
            raise ValueError(f"parameter_sets must have shape ({self.members}, {circuit.n_parameters}), got "
                             f"{params.shape}")
        self._sim.propagate(majoranas, mapping, gen_coeffs, params)

    def expectation_values(self) -> np.ndarray:
        """``<ref| O_q(theta_q) |ref>`` per member."""
        return self._sim.expectation_values()

    def sizes(self) -> np.ndarray:
        """Stored terms per member."""
        return self._sim.sizes()

    def size(self) -> int:
        return self._sim.size()

    def member_rows(self, q: int):
        """Member ``q``'s store ``(words[n, nw], coeffs[n])`` in its insertion order."""
        w, c, m = self._sim.rows()
        sel = m == q
        return w[sel], c[sel]

    def last_propagate_ms(self) -> float:
        return self._sim.last_propagate_ms()

    def gate_stats(self):
        return self._sim.gate_stats()


def _out_of_memory(err: RuntimeError) -> bool:
    msg = str(err)
    return any(s in msg for s in ("out of device memory", "out of memory", "32-bit row index"))


def expectation_grid(circuit: Circuit, operators: Sequence[Operator], parameter_sets, *, cutoff: int,
                     lower_atol: float | None = None, upper_atol: float | None = None,
                     cutoff_type: str | None = None, max_members: int = BatchSimulator.max_members,
                     stats: list | None = None) -> np.ndarray:
    """``<O_j>(theta_i)`` for every parameter set ``i`` and observable ``j``: shape
    ``(len(parameter_sets), len(operators))``.

    Members are the pairs in row-major order, taken ``max_members`` at a time; a batch whose store
    outgrows device memory is split in halves and run again.  ``stats``, when given, receives one
    ``(members, stored terms, propagate ms)`` tuple per batch that completed.
    """
    params = np.asarray(parameter_sets, dtype=np.float64)
    if params.ndim != 2:
        raise ValueError("parameter_sets must be a 2-D array (sets x circuit parameters)")
    n_sets, n_ops = params.shape[0], len(operators)
    pairs = [(i, j) for i in range(n_sets) for j in range(n_ops)]
    out = np.empty((n_sets, n_ops), dtype=np.float64)
    todo = [pairs[k:k + max_members] for k in range(0, len(pairs), max_members)]
    state = list(circuit.initial_state)
    while todo:
        chunk = todo.pop(0)
        b = None
        try:
            b = BatchPropagator([operators[j] for _, j in chunk], state, cutoff=cutoff, lower_atol=lower_atol,
                                upper_atol=upper_atol, cutoff_type=cutoff_type)
            b.propagate(circuit, params[[i for i, _ in chunk]])
            vals = b.expectation_values()
            if stats is not None:
                stats.append((len(chunk), b.size(), b.last_propagate_ms()))
        except RuntimeError as err:
            b = None  # give the store back before the halves run
            if len(chunk) == 1 or not _out_of_memory(err):
                raise
            half = len(chunk) // 2
            todo[:0] = [chunk[:half], chunk[half:]]
            continue
        finally:
            b = None
        for (i, j), v in zip(chunk, vals):
            out[i, j] = v
    return out
