"""only_rotate_len_k through the public API, with the reference's own cases
(tests/test_only_rotate_k.py:48-275): None is the uncapped default, a capped Schrodinger
rotation leaves the long terms untouched and rotates the short ones as the uncapped run does, the
S0_8e8o circuit with its orbital tail capped at 4 (graph and in-place) stores fewer terms and records
and keeps the energy, the capped graph build matches the oracle's rows, replay and gradient, and
out-of-range caps raise before anything runs."""

from __future__ import annotations

from contextlib import nullcontext

import numpy as np
import pytest

import fixtures

from monoprop_amd._engine import DeviceSimulator
from monoprop_amd.circuit import expand_monomials
from monoprop_amd import (Circuit, ExpGate, MajoranaOperator, MajoranaPropagator, PauliOperator,
                          PauliPropagator)

pytestmark = pytest.mark.gpu

KW = {"cutoff": 6, "schrodinger_cutoff": 8}


@pytest.mark.parametrize(("cls", "operator", "generator"), [
    (MajoranaPropagator, MajoranaOperator({(0,): 1.0}, 1), MajoranaOperator({(0, 1): 1.0j}, 1)),
    (PauliPropagator, PauliOperator({"Z": 1.0}, num_qubits=1), PauliOperator({"X": 1.0}, num_qubits=1)),
])
@pytest.mark.parametrize(("builder", "result"), [("build_graph", "expval_functional"), ("propagate", "expval")])
def test_none_is_uncapped(cls, operator, generator, builder, result):
    print(f"case: only_rotate None {cls.__name__} {builder}", flush=True)
    circuit = Circuit((ExpGate(generator, index=0),), initial_state=(), system_size=1, parameters=(0.321,))

    def ev(mp):
        r = getattr(mp, result)
        return r()(circuit.parameters) if result == "expval_functional" else r()

    a = cls(operator, circuit.initial_state, **KW)
    getattr(a, builder)(circuit)
    b = cls(operator, circuit.initial_state, **KW)
    getattr(b, builder)(circuit, only_rotate_len_k=None)
    assert a.graph_size() == b.graph_size() and a.size() == b.size()
    assert np.isclose(ev(a), ev(b), atol=1e-12)


def test_basic_orbital_rotation():
    print("case: only_rotate orbital", flush=True)
    n = 4
    op = MajoranaOperator({}, n)
    circuit = Circuit.from_dense_arrays(initial_state=[], majoranas=[(1, 2)], parameters=[np.pi / 4], gen_coeffs=[1.0],
                                        param_inds=[0], system_size=n)
    act = MajoranaPropagator(op, circuit.initial_state, **KW)
    initial = act.evolved_operator()
    act.propagate(circuit)
    rotated = act.evolved_operator()
    orb = MajoranaPropagator(op, circuit.initial_state, **KW)
    orb.propagate(circuit, only_rotate_len_k=4)
    orb_rotated = orb.evolved_operator()
    long_keys = [k for k in initial.terms if len(k) > 4]
    short_keys = [k for k in rotated.terms if len(k) <= 4]
    assert long_keys and short_keys
    for k in long_keys:
        assert np.isclose(initial.terms[k], orb_rotated.terms[k], atol=1e-12)
    for k in short_keys:
        assert np.isclose(rotated.terms[k], orb_rotated.terms[k], atol=1e-12)


CASES = [(-1, True), (0, True), (None, False), (9, True), (8, False)]


@pytest.mark.parametrize(("k", "raises"), CASES)
@pytest.mark.parametrize("method", ["build_graph", "propagate"])
@pytest.mark.parametrize("basis", ["majorana", "pauli"])
def test_out_of_range_caps_raise(k, raises, method, basis):
    mp = (MajoranaPropagator(MajoranaOperator({}, 4), [], **KW) if basis == "majorana"
          else PauliPropagator(PauliOperator({}, 4), [], **KW))
    ctx = (pytest.raises(ValueError, match=rf"only_rotate_len_k={k} is out of range; must be 0 < k <= 2\*num_qubits")
           if raises else nullcontext())
    with ctx:
        getattr(mp, method)(Circuit((), initial_state=(), system_size=4), only_rotate_len_k=k)


def _split_orbital_gates(gates):
    """(non-orbital, orbital): the orbital gates are the tail run whose monomials all have length 2."""
    def is_orbital(gate):
        return all(len(m) == 2 for m in gate.generator.terms)

    for i in range(len(gates)):
        if all(is_orbital(g) for g in gates[i:]):
            return gates[:i], gates[i:]
    return gates, []


@pytest.mark.parametrize("inplace", [False, True])
def test_only_rotate_len_k_fixture(inplace):
    """The reference's test_only_rotate_len_k (tests/test_only_rotate_k.py:149-190) on S0_8e8o_c8."""
    print(f"case: only_rotate S0_8e8o orbital tail inplace={inplace}", flush=True)
    p = fixtures.load("S0_8e8o_majoranic_c8")
    circuit = p.circuit()
    parameters = circuit.parameters
    non_orbital_gates, orbital_gates = _split_orbital_gates(list(circuit.gates))
    assert orbital_gates
    split = len(non_orbital_gates)
    non_orbital = Circuit(tuple(ExpGate._with_index(g, None) for g in non_orbital_gates), system_size=p.n_modes,
                          parameters=tuple(parameters[:split]))
    orbital = Circuit(tuple(ExpGate._with_index(g, None) for g in orbital_gates), system_size=p.n_modes,
                      parameters=tuple(parameters[split:]))
    act = MajoranaPropagator(p.operator, circuit.initial_state, **KW)
    act.build_graph(circuit)
    act_ener = act.expval_functional()(parameters)
    mp = MajoranaPropagator(p.operator, circuit.initial_state, **KW)
    if inplace:
        mp.propagate(non_orbital)
        mp.propagate(orbital, only_rotate_len_k=4)
        test_expval = mp.expval()
    else:
        mp.build_graph(non_orbital)
        mp.build_graph(orbital, only_rotate_len_k=4)
        test_expval = mp.expval_functional()(parameters)
        assert sum(mp.graph_size()) < sum(act.graph_size())
    assert mp.size() < act.size()
    assert np.isclose(test_expval, act_ener, atol=1e-12)


@pytest.mark.parametrize("k", [2, 4])
@pytest.mark.parametrize("seeded", [False, True])
def test_capped_graph_matches_oracle(oracle_mod, k, seeded):
    """A capped graph build of the whole S0_8e8o circuit at generic angles: the store, the replayed
    coefficient vector and <O> are bit-exact with the oracle, the adjoint gradient within 1e-10."""
    print(f"case: capped graph S0_8e8o k={k} seeded={seeded}", flush=True)
    p = fixtures.load("S0_8e8o_majoranic_c8")
    circ = p.circuit()
    n = p.n_modes
    monos, gcs, pm, gidx = expand_monomials(circ.gates, circ.resolved_mapping, n)
    x = np.random.default_rng(3).uniform(-1.0, 1.0, len(circ.parameters))
    seed = list(x) if seeded else None
    o = oracle_mod.Oracle(p.operator.terms, n, basis="majorana", cutoff=8, cutoff_type="length", lower_atol=1e-8,
                          initial_state=p.initial_state)
    o.build_graph(monos, pm, gcs, seed, only_rotate_len_k=k)
    d = DeviceSimulator(p.operator.terms, 8, list(p.initial_state), lower_atol=1e-8, cutoff_type="length",
                        logical_num_modes=n, basis="majorana")
    d.build_graph(monos, pm, gcs, gidx, seed, only_rotate_len_k=k)
    wo, _ = o.rows()
    wd, _ = d.rows()
    assert np.array_equal(wo, wd)
    assert np.array_equal(np.asarray(d.contract_partially(x, False)), o.graph_contract(x))
    vo, go = o.graph_grad(x)
    vd, gd = d.expectation_value_and_gradient(x)
    assert vo == vd == d.expectation_value(x)
    assert np.allclose(gd, go, rtol=1e-10, atol=1e-12)
