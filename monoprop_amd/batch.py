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
        self._config = dict(cutoff=cutoff, lower_atol=lower_atol, upper_atol=upper_atol, cutoff_type=cutoff_type)
        # identity coefficients (the engine's core terms): the real part, as both encodings keep it
        self._cores = np.asarray([complex(t.get((), 0.0)).real for _, t, _ in encoded], dtype=np.float64)
        self._sim = BatchSimulator([t for _, t, _ in encoded], cutoff, self._initial_state,
                                   logical_num_modes=self._num_modes, basis=basis, **self._drop_cutoff())

    def _drop_cutoff(self) -> dict:
        return {k: v for k, v in self._config.items() if k != "cutoff"}

    @property
    def _spec(self):
        return self._basis, self._num_modes, self._initial_state, dict(self._config)

    @classmethod
    def _from_rows(cls, spec, cores, words, coeffs, members) -> BatchPropagator:
        """A batch of configuration ``spec`` holding the given rows (members numbered 0..)."""
        b = cls.__new__(cls)
        b._basis, b._num_modes, b._initial_state, b._config = spec
        b._cores = np.asarray(cores, dtype=np.float64)
        b._sim = BatchSimulator([], b._config["cutoff"], b._initial_state, logical_num_modes=b._num_modes,
                                basis=b._basis, _rows=(b._cores, words, coeffs, members), **b._drop_cutoff())
        return b

    @property
    def members(self) -> int:
        return self._sim.members

    @property
    def num_modes(self) -> int:
        return self._num_modes

    def _layers(self, circuit: Circuit):
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
        return majoranas, mapping, gen_coeffs

    def propagate(self, circuit: Circuit, parameter_sets) -> None:
        """Evolve every member through ``circuit`` (Heisenberg, back to front) at its own row of
        ``parameter_sets`` (shape ``(members, circuit.n_parameters)``)."""
        majoranas, mapping, gen_coeffs = self._layers(circuit)
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

    def split(self, move) -> BatchPropagator:
        """Move the members with ``move[q]`` true into a new batch on the device (store order kept,
        members renumbered in ascending order on both sides); both continue from where this one
        stopped."""
        mv = np.asarray(move, dtype=bool)
        b = BatchPropagator.__new__(BatchPropagator)
        b._basis, b._num_modes, b._initial_state, b._config = self._spec
        b._sim = self._sim.split(mv)
        b._cores = self._cores[mv]
        self._cores = self._cores[~mv]
        return b

    def member_rows(self, q: int):
        """Member ``q``'s store ``(words[n, nw], coeffs[n])`` in its insertion order."""
        w, c, m = self._sim.rows()
        sel = m == q
        return w[sel], c[sel]

    def last_propagate_ms(self) -> float:
        return self._sim.last_propagate_ms()

    def gate_stats(self):
        return self._sim.gate_stats()


def _halves(sizes: np.ndarray) -> np.ndarray:
    """Two groups of members with near-equal stored terms (largest first, each to the lighter group);
    True marks the second group.  Both groups are non-empty for two or more members."""
    side = np.zeros(len(sizes), dtype=bool)
    load = [0, 0]
    for q in np.argsort(-np.asarray(sizes), kind="stable"):
        g = 0 if load[0] <= load[1] else 1
        side[q] = g == 1
        load[g] += int(sizes[q]) + 1
    return side


def expectation_grid(circuit: Circuit, operators: Sequence[Operator], parameter_sets, *, cutoff: int,
                     lower_atol: float | None = None, upper_atol: float | None = None,
                     cutoff_type: str | None = None, max_members: int = BatchSimulator.max_members,
                     max_capacity: int = -1, stats: list | None = None, progress=None,
                     workers: int = 1) -> np.ndarray:
    """``<O_j>(theta_i)`` for every parameter set ``i`` and observable ``j``: shape
    ``(len(parameter_sets), len(operators))``.

    Members are the pairs in row-major order, ``max_members`` per batch.  A batch of several members
    whose store would grow past ``max_capacity`` rows (< 0: sized to the device, 0: no bound) stops
    before that gate; its members are split into two halves of near-equal size, which continue from
    that gate as separate batches (depth first; the waiting half stays on the device at its exact
    size).  A single
    member is never bounded.  ``stats``, when given, receives one ``(members, stored terms, device
    ms, split, term-gate applications)`` tuple per propagate call, ``split`` true when the call
    stopped at the bound; ``progress``, when given, is called with each such tuple.

    ``workers`` > 1 cuts the parameter sets into that many contiguous groups and runs one driver per
    group at once (``monoprop_amd.concurrent``: own engines, own streams, own host threads); the
    automatic bound then shares the available memory among them.  Every value is the same to the bit
    whatever the grouping (a member's propagation does not depend on its batch).
    """
    params = np.asarray(parameter_sets, dtype=np.float64)
    if params.ndim != 2:
        raise ValueError("parameter_sets must be a 2-D array (sets x circuit parameters)")
    if workers < 1:
        raise ValueError("workers must be at least 1")
    if workers > 1 and params.shape[0] > 1:
        from .concurrent import run_concurrently

        groups = [g for g in np.array_split(np.arange(params.shape[0]), min(workers, params.shape[0])) if len(g)]
        cap = -len(groups) if max_capacity < 0 else max_capacity
        parts = run_concurrently([
            lambda g=g: expectation_grid(circuit, operators, params[g], cutoff=cutoff, lower_atol=lower_atol,
                                         upper_atol=upper_atol, cutoff_type=cutoff_type, max_members=max_members,
                                         max_capacity=cap, stats=stats, progress=progress)
            for g in groups])
        return np.concatenate(parts, axis=0)
    n_sets, n_ops = params.shape[0], len(operators)
    pairs = [(i, j) for i in range(n_sets) for j in range(n_ops)]
    out = np.empty((n_sets, n_ops), dtype=np.float64)
    state = list(circuit.initial_state)
    # job: (pairs, layers still to apply, the stopped batch holding them or None)
    jobs = [(pairs[k:k + max_members], None, None) for k in range(0, len(pairs), max_members)][::-1]
    layers = None
    while jobs:
        chunk, left, b = jobs.pop()
        if b is None:
            b = BatchPropagator([operators[j] for _, j in chunk], state, cutoff=cutoff, lower_atol=lower_atol,
                                upper_atol=upper_atol, cutoff_type=cutoff_type)
        if layers is None:
            layers = b._layers(circuit)
            if params.shape[1] != circuit.n_parameters:
                raise ValueError(f"parameter_sets must have {circuit.n_parameters} columns, got {params.shape[1]}")
        majoranas, mapping, gen_coeffs = layers
        if left is None:
            left = len(majoranas)
        bound = max_capacity if len(chunk) > 1 else 0
        done = b._sim.propagate(majoranas[:left], mapping[:left], gen_coeffs[:left], params[[i for i, _ in chunk]],
                                bound)
        left -= done
        entry = (len(chunk), b.size(), b.last_propagate_ms(), left > 0, int(b.gate_stats()[0].sum()))
        if stats is not None:
            stats.append(entry)
        if progress is not None:
            progress(entry)
        if left == 0:
            for (i, j), v in zip(chunk, b.expectation_values()):
                out[i, j] = v
            continue
        move = _halves(b.sizes())
        moved = b.split(move)
        jobs.append(([p for p, f in zip(chunk, move) if f], left, moved))
        jobs.append(([p for p, f in zip(chunk, move) if not f], left, b))
        b = moved = None
    return out
