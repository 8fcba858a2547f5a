"""Host logic of the batch sweep driver (no GPU): batching order, halving on device-memory
exhaustion, and that other failures propagate."""

from __future__ import annotations

import numpy as np
import pytest

from monoprop_amd import batch
from monoprop_amd.pauli import Pauli, PauliOperator


class _FakeCircuit:
    initial_state = ()


class _FakeBatch:
    """Stands in for the device store: <O> of member (set i, observable j) = 10 i + j; batches above
    `limit` members fail the way a store that outgrows device memory does."""

    limit = 4
    fail_with = "out of device memory growing the term store to 123 rows (61 live terms)"
    created: list[int] = []

    def __init__(self, ops, state, **kw):
        self.ops = ops
        _FakeBatch.created.append(len(ops))

    def propagate(self, circuit, params):
        if len(self.ops) > self.limit:
            raise RuntimeError(self.fail_with)
        self.vals = [10 * p[0] + op for p, op in zip(params, self.ops)]

    def expectation_values(self):
        return np.asarray(self.vals, dtype=float)

    def size(self):
        return 0

    def last_propagate_ms(self):
        return 0.0


@pytest.fixture
def fake(monkeypatch):
    monkeypatch.setattr(batch, "BatchPropagator", _FakeBatch)
    _FakeBatch.created = []
    _FakeBatch.limit = 4
    _FakeBatch.fail_with = "out of device memory growing the term store to 123 rows (61 live terms)"
    return _FakeBatch


def test_grid_values_in_set_major_order(fake):
    params = np.arange(6, dtype=float).reshape(6, 1)
    grid = batch.expectation_grid(_FakeCircuit(), [0, 1, 2], params, cutoff=4, max_members=4)
    assert grid.shape == (6, 3)
    assert np.array_equal(grid, 10 * np.arange(6)[:, None] + np.arange(3)[None, :])
    assert fake.created == [4, 4, 4, 4, 2]


def test_grid_halves_batches_that_run_out_of_memory(fake):
    params = np.arange(5, dtype=float).reshape(5, 1)
    fake.limit = 3
    grid = batch.expectation_grid(_FakeCircuit(), [0, 1], params, cutoff=4, max_members=8)
    assert np.array_equal(grid, 10 * np.arange(5)[:, None] + np.arange(2)[None, :])
    # 8 fails -> 4 + 4; each 4 fails -> 2 + 2; then the last batch of 2
    assert fake.created == [8, 4, 2, 2, 4, 2, 2, 2]


def test_grid_reraises_other_failures_and_single_members(fake):
    params = np.zeros((2, 1))
    fake.fail_with = "device pipeline wait limit exceeded"
    fake.limit = 1
    with pytest.raises(RuntimeError, match="wait limit"):
        batch.expectation_grid(_FakeCircuit(), [0], params, cutoff=4)
    fake.fail_with = "out of device memory"
    fake.limit = 0
    with pytest.raises(RuntimeError, match="out of device memory"):
        batch.expectation_grid(_FakeCircuit(), [0], params, cutoff=4)


def test_local_terms_encoding():
    basis, terms, n = batch._local_terms(PauliOperator({Pauli("ZX", (1, 2)): 0.5}, 4))
    assert basis == "pauli" and n == 4
    assert terms == {(2, 3, 4): 0.5}
    with pytest.raises(TypeError):
        batch._local_terms({(0, 1): 1.0})
