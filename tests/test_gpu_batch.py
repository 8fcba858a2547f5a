"""Batched propagation vs the oracle.

Every member of a batch store must hold, in store order, exactly the rows and FP64 coefficients the
oracle produces propagating that member alone (bit-exact), the same per-member term count, and <O>
to the bit.  Members share the circuit structure and differ in observable and angles; identical
members and identity-only members check that equal monomials of different members never merge.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from monoprop_amd import workloads
from monoprop_amd._engine import load_library
from monoprop_amd.batch import BatchPropagator, expectation_grid
from monoprop_amd.circuit import expand_monomials
from monoprop_amd.majorana import MajoranaOperator
from monoprop_amd.pauli import Pauli, PauliOperator

pytestmark = pytest.mark.gpu


def _oracle(oracle_mod, basis, op, circ, params, *, cutoff, cutoff_type, lower=None, upper=None):
    n = circ.system_size
    terms = op.get_local_operator().terms if basis == "pauli" else op.terms
    monos, gcs, pm, _ = expand_monomials(circ.gates, circ.resolved_mapping, n)
    o = oracle_mod.Oracle(terms, n, basis=basis, cutoff=cutoff, cutoff_type=cutoff_type, lower_atol=lower,
                          upper_atol=upper, initial_state=list(circ.initial_state))
    o.propagate(monos, pm, gcs, list(params))
    return o


def _check_members(oracle_mod, b, basis, ops, circ, params, **trunc):
    w, c, m = b._sim.rows()
    sizes = b.sizes()
    ev = b.expectation_values()
    assert int(sizes.sum()) == b.size() == len(m)
    for q, op in enumerate(ops):
        o = _oracle(oracle_mod, basis, op, circ, params[q], **trunc)
        wo, co = o.rows()
        sel = m == q
        assert np.array_equal(w[sel], wo), f"member {q}: rows differ"
        assert np.array_equal(c[sel], co), f"member {q}: coefficients differ"
        assert sizes[q] == o.size()
        assert ev[q] == o.expectation(), (q, ev[q], o.expectation())


def _chain_params(circ, sets):
    """Per member one (theta_J, theta_h) pair mapped onto the circuit's ZZ / X gates."""
    return workloads.kicked_ising_rows(circ, sets)


def test_batch_chain_members(oracle_mod):
    """20-qubit chain, 3 steps: 4 angle sets x 3 observables as 12 members of one store."""
    print("case: batch chain n=20 steps=3 members=12", flush=True)
    n = 20
    circ = workloads.kicked_ising_chain(n, 3)
    rng = np.random.default_rng(5)
    sets = rng.uniform(-np.pi / 2, np.pi / 2, size=(4, 2))
    obs = [PauliOperator({Pauli("Z", (10,)): 1.0}, n),
           PauliOperator({Pauli("ZZ", (4, 5)): 0.5, Pauli("X", (12,)): -0.25}, n),
           PauliOperator({Pauli("XXXXX", (2, 3, 4, 5, 6)): 1.0}, n)]
    ops = [obs[j] for i in range(4) for j in range(3)]
    params = _chain_params(circ, [sets[i] for i in range(4) for _ in range(3)])
    b = BatchPropagator(ops, [], cutoff=8, lower_atol=1e-8)
    b.propagate(circ, params)
    assert b.size() > 1000
    _check_members(oracle_mod, b, "pauli", ops, circ, params, cutoff=8, cutoff_type="support", lower=1e-8)


def test_batch_identical_and_identity_members(oracle_mod):
    """Two identical members stay separate; an identity-only member keeps an empty store and <O> = core."""
    print("case: batch identical members", flush=True)
    n = 12
    circ = workloads.kicked_ising_chain(n, 2)
    z = PauliOperator({Pauli("Z", (6,)): 1.0}, n)
    ident = PauliOperator({Pauli("", ()): 0.75}, n)
    ops = [z, z, ident, z]
    params = _chain_params(circ, [(0.3, -0.7), (0.3, -0.7), (0.1, 0.2), (-1.1, 0.4)])
    b = BatchPropagator(ops, [], cutoff=6, lower_atol=1e-8)
    b.propagate(circ, params)
    sizes = b.sizes()
    assert sizes[0] == sizes[1] > 0 and sizes[2] == 0
    ev = b.expectation_values()
    assert ev[0] == ev[1] and ev[2] == 0.75
    _check_members(oracle_mod, b, "pauli", ops, circ, params, cutoff=6, cutoff_type="support", lower=1e-8)


def test_batch_initial_state_and_upper_atol(oracle_mod):
    """A non-trivial reference state (phases in <O>) and the upper_atol rescue."""
    print("case: batch state + upper", flush=True)
    n = 16
    circ = workloads.kicked_ising_chain(n, 2)
    circ = type(circ)(circ.gates, n, circ.parameters, initial_state=[0, 3, 8])
    ops = [PauliOperator({Pauli("ZZ", (3, 4)): 1.0}, n), PauliOperator({Pauli("Z", (8,)): 1.0}, n)]
    params = _chain_params(circ, [(0.9, 0.35), (-0.4, 1.2)])
    b = BatchPropagator(ops, [0, 3, 8], cutoff=3, lower_atol=1e-4, upper_atol=5e-2)
    b.propagate(circ, params)
    _check_members(oracle_mod, b, "pauli", ops, circ, params, cutoff=3, cutoff_type="support", lower=1e-4,
                   upper=5e-2)


@pytest.mark.parametrize("cutoff_type", ["length", "support"])
def test_batch_majorana_hubbard(oracle_mod, cutoff_type):
    """3x3 Hubbard, 2 steps: number operators at different modes, different (t, U) angles."""
    print(f"case: batch hubbard 3x3 {cutoff_type}", flush=True)
    circ, n = workloads.hubbard_lattice(3, 2)
    ops = [workloads.number_operator(k, n) for k in (0, 3, 7, 0)]
    rng = np.random.default_rng(11)
    params = np.tile(np.asarray(circ.parameters), (4, 1)) * rng.uniform(0.5, 2.0, size=(4, 1))
    b = BatchPropagator(ops, list(circ.initial_state), cutoff=6, lower_atol=1e-6, cutoff_type=cutoff_type)
    b.propagate(circ, params)
    _check_members(oracle_mod, b, "majorana", ops, circ, params, cutoff=6, cutoff_type=cutoff_type, lower=1e-6)


def test_batch_lattice_regrowth(oracle_mod):
    """10x10 lattice, 3 steps (840 gates: several enqueue chunks and store regrowths), 3 members."""
    print("case: batch lattice 10x10 steps=3", flush=True)
    circ = workloads.kicked_ising_lattice(10, 3)
    n = 100
    ops = [PauliOperator({Pauli("Z", (s,)): 0.01 for s in range(n)}, n),
           PauliOperator({Pauli("ZZZZ", (44, 45, 54, 55)): 1.0}, n),
           PauliOperator({Pauli("Z", (63,)): 1.0, Pauli("Z", (64,)): 1.0}, n)]
    params = _chain_params(circ, [(-np.pi / 4, 0.8), (0.6, 0.5), (-0.3, 1.3)])
    b = BatchPropagator(ops, [], cutoff=10, lower_atol=1e-5)
    b.propagate(circ, params)
    _check_members(oracle_mod, b, "pauli", ops, circ, params, cutoff=10, cutoff_type="support", lower=1e-5)


def test_expectation_grid_small_batches(oracle_mod):
    """The sweep driver over 5 sets x 2 observables in batches of at most 3 members."""
    print("case: expectation grid", flush=True)
    n = 14
    circ = workloads.kicked_ising_chain(n, 2)
    rng = np.random.default_rng(3)
    sets = rng.uniform(-np.pi / 2, np.pi / 2, size=(5, 2))
    params = _chain_params(circ, sets)
    obs = [PauliOperator({Pauli("Z", (7,)): 1.0}, n), PauliOperator({Pauli("ZZ", (2, 3)): 1.0}, n)]
    stats = []
    grid = expectation_grid(circ, obs, params, cutoff=6, lower_atol=1e-6, max_members=3, stats=stats)
    assert [s[0] for s in stats] == [3, 3, 3, 1]
    for i in range(5):
        for j in range(2):
            o = _oracle(oracle_mod, "pauli", obs[j], circ, params[i], cutoff=6, cutoff_type="support", lower=1e-6)
            assert grid[i, j] == o.expectation()


def test_batch_refuses_single_entry_points():
    """Entry points of a lone propagator return -1 on a batch engine; bad parameter shapes are ValueError."""
    n = 8
    circ = workloads.kicked_ising_chain(n, 1)
    b = BatchPropagator([PauliOperator({Pauli("Z", (3,)): 1.0}, n)] * 2, [], cutoff=4)
    with pytest.raises(ValueError):
        b.propagate(circ, np.zeros((3, circ.n_parameters)))
    lib = load_library()
    out = C.c_double()
    assert lib.mpa_expectation_value(b._sim._h, None, 0, C.byref(out)) == -1
    assert b"batch" in lib.mpa_last_error()
    with pytest.raises(ValueError):
        BatchPropagator([PauliOperator({Pauli("Z", (0,)): 1.0}, 8), MajoranaOperator({(0, 1): 1j}, 8)], [], cutoff=4)


def _c1_members(n_sets=3):
    n = 20
    circ = workloads.kicked_ising_chain(n, 5)
    rng = np.random.default_rng(17)
    sets = rng.uniform(-np.pi / 2, np.pi / 2, size=(n_sets, 2))
    obs = [PauliOperator({Pauli("Z", (10,)): 1.0}, n), PauliOperator({Pauli("ZZ", (9, 10)): 1.0}, n)]
    return circ, sets, obs


def test_stop_and_resume_is_bit_exact(oracle_mod):
    """A batch stopped by its capacity bound and rebuilt from its rows continues exactly where it stopped:
    every member's store equals the oracle's."""
    print("case: batch stop + resume c1", flush=True)
    circ, sets, obs = _c1_members()
    ops = [obs[j] for i in range(len(sets)) for j in range(len(obs))]
    params = _chain_params(circ, [sets[i] for i in range(len(sets)) for _ in obs])
    b = BatchPropagator(ops, [], cutoff=8, lower_atol=1e-8)
    majoranas, mapping, gen_coeffs = b._layers(circ)
    done = b._sim.propagate(majoranas, mapping, gen_coeffs, params, 1 << 15)
    assert 0 < done < len(majoranas)
    w, c, m = b._sim.rows()
    b2 = BatchPropagator._from_rows(b._spec, b._cores, w, c, m)
    del b
    left = len(majoranas) - done
    assert b2._sim.propagate(majoranas[:left], mapping[:left], gen_coeffs[:left], params, 0) == left
    _check_members(oracle_mod, b2, "pauli", ops, circ, params, cutoff=8, cutoff_type="support", lower=1e-8)


def test_expectation_grid_splits_match_oracle(oracle_mod):
    """A capacity bound of 2^15 rows makes the sweep stop, split the members into halves and resume them;
    every value still matches the oracle's."""
    print("case: expectation grid with splits c1", flush=True)
    circ, sets, obs = _c1_members()
    params = _chain_params(circ, sets)
    stats = []
    grid = expectation_grid(circ, obs, params, cutoff=8, lower_atol=1e-8, max_capacity=1 << 15, stats=stats)
    assert any(s[3] for s in stats), stats
    for i in range(len(sets)):
        for j in range(len(obs)):
            o = _oracle(oracle_mod, "pauli", obs[j], circ, params[i], cutoff=8, cutoff_type="support", lower=1e-8)
            assert grid[i, j] == o.expectation(), (i, j)


def test_device_split_then_resume_is_bit_exact(oracle_mod):
    """A batch stopped by its bound, split on the device into two batches (members 1, 2, 5 move), and both
    continued: every member's store equals the oracle's, members renumbered in ascending order."""
    print("case: batch device split c1", flush=True)
    circ, sets, obs = _c1_members()
    ops = [obs[j] for i in range(len(sets)) for j in range(len(obs))]
    params = _chain_params(circ, [sets[i] for i in range(len(sets)) for _ in obs])
    b = BatchPropagator(ops, [], cutoff=8, lower_atol=1e-8)
    majoranas, mapping, gen_coeffs = b._layers(circ)
    done = b._sim.propagate(majoranas, mapping, gen_coeffs, params, 1 << 15)
    assert 0 < done < len(majoranas)
    before = b.sizes()
    move = np.zeros(len(ops), dtype=bool)
    move[[1, 2, 5]] = True
    other = b.split(move)
    assert np.array_equal(other.sizes(), before[move]) and np.array_equal(b.sizes(), before[~move])
    left = len(majoranas) - done
    for part, sel in ((b, ~move), (other, move)):
        assert part._sim.propagate(majoranas[:left], mapping[:left], gen_coeffs[:left], params[sel], 0) == left
        _check_members(oracle_mod, part, "pauli", [o for o, f in zip(ops, sel) if f], circ, params[sel], cutoff=8,
                       cutoff_type="support", lower=1e-8)
