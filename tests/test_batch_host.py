"""Host logic of the batch sweep driver (no GPU): batching order, the split of a batch that reaches
its capacity bound (each half continues from the gate where the batch stopped, so no layer is
applied twice or skipped), and the size-balanced halves."""

from __future__ import annotations

import numpy as np
import pytest

from monoprop_amd import batch
from monoprop_amd.pauli import Pauli, PauliOperator

N_LAYERS = 10


class _FakeCircuit:
    initial_state = ()
    n_parameters = 1


class _FakeSim:
    """One row per member: words = (layers applied, observable j), coeff = accumulated value.  Member
    q grows by (j + 1) rows per layer; a layer whose growth would take the total past the bound is
    not applied (the stop).  Applying a layer at set i adds 10 i + j to the member's value."""

    def __init__(self, words, coeffs, members):
        self.w = np.array(words, dtype=np.uint64).reshape(-1, 2)
        self.c = np.array(coeffs, dtype=float)
        self.m = np.array(members, dtype=np.int32)
        self.calls = []

    def sizes(self):
        return (self.w[:, 0] * (self.w[:, 1] + 1)).astype(np.int64)

    def propagate(self, majoranas, mapping, gen, prm, bound):
        left = len(majoranas)
        grow = (self.w[:, 1] + 1).astype(int)
        k = 0
        while k < left and (bound == 0 or int(((self.w[:, 0] + k + 1) * grow).sum()) <= bound):
            k += 1
        weight = 10 * np.asarray(prm)[self.m, 0] + self.w[:, 1]
        self.w[:, 0] += k
        self.c += k * weight
        _FakeBatch.calls.append((len(self.c), left, k, bound))
        return k

    def rows(self):
        return self.w.copy(), self.c.copy(), self.m.copy()


class _FakeBatch:
    calls: list = []

    def __init__(self, ops, state, **kw):
        self._sim = _FakeSim([(0, j) for j in ops], np.zeros(len(ops)), np.arange(len(ops)))
        self._cores = np.zeros(len(ops))
        self._spec = "spec"

    def split(self, move):
        """Rows of the flagged members to a new batch, store order kept, both sides renumbered."""
        move = np.asarray(move, dtype=bool)
        w, c, m = self._sim.rows()
        parts = []
        for flag in (True, False):
            keep = move == flag
            renum = np.cumsum(keep) - 1
            sel = keep[m]
            parts.append(_FakeSim(w[sel], c[sel], renum[m[sel]]))
        other = _FakeBatch.__new__(_FakeBatch)
        other._sim, self._sim = parts
        other._cores, self._cores = self._cores[move], self._cores[~move]
        return other

    def _layers(self, circuit):
        return [()] * N_LAYERS, [0] * N_LAYERS, [1.0] * N_LAYERS

    def sizes(self):
        return self._sim.sizes()

    def size(self):
        return int(self.sizes().sum())

    def expectation_values(self):
        return self._sim.c.copy()

    def last_propagate_ms(self):
        return 0.0

    def gate_stats(self):
        z = np.zeros(N_LAYERS, dtype=np.int64)
        return z, z, z, z


@pytest.fixture
def fake(monkeypatch):
    monkeypatch.setattr(batch, "BatchPropagator", _FakeBatch)
    _FakeBatch.calls = []
    return _FakeBatch


def _expected(n_sets, ops):
    return N_LAYERS * (10 * np.arange(n_sets)[:, None] + np.asarray(ops)[None, :])


def test_grid_batches_in_set_major_order(fake):
    params = np.arange(6, dtype=float).reshape(6, 1)
    grid = batch.expectation_grid(_FakeCircuit(), [0, 1, 2], params, cutoff=4, max_members=4, max_capacity=0)
    assert np.array_equal(grid, _expected(6, [0, 1, 2]))
    assert [c[0] for c in fake.calls] == [4, 4, 4, 4, 2]
    assert all(c[2] == N_LAYERS for c in fake.calls)


@pytest.mark.parametrize("bound,first_done", [(100, 3), (25, 0)])
def test_grid_splits_at_the_bound_and_resumes(fake, bound, first_done):
    """15 members growing 30 rows per layer: the batch stops where the next layer would pass the bound
    and halves recursively; every member still gets each of the 10 layers exactly once."""
    params = np.arange(5, dtype=float).reshape(5, 1)
    stats = []
    grid = batch.expectation_grid(_FakeCircuit(), [0, 1, 2], params, cutoff=4, max_capacity=bound, stats=stats)
    assert np.array_equal(grid, _expected(5, [0, 1, 2]))
    assert fake.calls[0] == (15, N_LAYERS, first_done, bound)
    assert fake.calls[1][1] == N_LAYERS - first_done  # the halves resume where the batch stopped
    assert any(s[3] for s in stats) and not stats[-1][3]
    assert sum(c[2] * c[0] for c in fake.calls) >= 15 * N_LAYERS
    if bound == 25:  # a member alone outgrows 25 rows: it runs unbounded
        singles = [c for c in fake.calls if c[0] == 1]
        assert singles and all(c[3] == 0 for c in singles)


def test_halves_balance_sizes():
    side = batch._halves(np.array([5, 100, 3, 50, 50, 0]))
    sizes = np.array([5, 100, 3, 50, 50, 0])
    assert side.any() and (~side).any()
    assert abs(int(sizes[side].sum()) - int(sizes[~side].sum())) <= 5


def test_local_terms_encoding():
    basis, terms, n = batch._local_terms(PauliOperator({Pauli("ZX", (1, 2)): 0.5}, 4))
    assert basis == "pauli" and n == 4
    assert terms == {(2, 3, 4): 0.5}
    with pytest.raises(TypeError):
        batch._local_terms({(0, 1): 1.0})
