"""Propagator-level circuit behaviour with the cases of the reference's tests/test_circuit.py:523-1022:
the parameter mapping getter/setter, n_gates, initial-state checks, the graph's parameter axis across
failed builds and in-place contractions, and propagate on top of a graph."""

from __future__ import annotations

import numpy as np
import pytest

import fixtures
from monoprop_amd import Circuit, ExpGate, MajoranaOperator, MajoranaPropagator, PauliOperator

pytestmark = pytest.mark.gpu

_OBS = MajoranaOperator({(0, 1, 2, 4): 1.0}, 8)


def _prop(p):
    return MajoranaPropagator(p.operator, p.initial_state, cutoff=2 * p.n_modes)


def _small(**kw):
    return MajoranaPropagator(_OBS, initial_state=[], cutoff=16, **kw)


def _two_gates(**kw):
    return Circuit((ExpGate(MajoranaOperator({(4, 5): -1.0j}, 8)), ExpGate(MajoranaOperator({(2, 3): -1.0j}, 8))),
                   system_size=8, parameters=(0.5, 0.3), **kw)


def test_mapping_getter_setter_on_fixture():
    print("case: circuit mapping", flush=True)
    p = fixtures.load("lih_fermionic_spin_exact")
    c = p.circuit()
    prop = _prop(p)
    prop.build_graph(c)
    m = prop.parameter_mapping
    assert len(m) == prop.graph_layers and max(m) + 1 == prop.n_parameters
    prop.parameter_mapping = m
    assert prop.parameter_mapping == m
    np.testing.assert_allclose(prop.expval(c), p.exact_energy)
    n = prop.graph_layers
    prop.parameter_mapping = [0] * n
    assert prop.n_parameters == 1
    tied = prop.expval([0.3])
    prop.parameter_mapping = list(range(n))
    assert prop.n_parameters == n
    np.testing.assert_allclose(tied, prop.expval([0.3] * n))


def test_mapping_setter_validates():
    p = fixtures.load("rx_rz_ry_rz_exact")
    prop = _prop(p)
    prop.build_graph(p.circuit())
    with pytest.raises(ValueError, match=r"per graph layer.*per gate"):
        prop.parameter_mapping = [0]
    with pytest.raises(ValueError, match="contiguous"):
        prop.parameter_mapping = [i + 1 for i in range(prop.graph_layers)]


def test_n_gates_and_per_gate_mapping():
    op = MajoranaOperator({(0, 1): 1.0j}, 2)
    prop = MajoranaPropagator(op, [0, 1], cutoff=4)
    g0 = ExpGate(MajoranaOperator({(0, 2): 1.0j, (1, 3): 1.0j}, 2))
    g1 = ExpGate(MajoranaOperator({(2,): 1.0}, 2))
    prop.build_graph(Circuit((g0, g1), initial_state=[0, 1], system_size=2))
    assert prop.n_gates == 2 and prop.graph_layers > prop.n_gates
    prop.parameter_mapping = [0, 0]
    assert prop.n_parameters == 1
    by_gate = prop.expval([0.3])
    prop.parameter_mapping = [0] * prop.graph_layers
    np.testing.assert_allclose(by_gate, prop.expval([0.3]))
    acc = MajoranaPropagator(op, [0, 1], cutoff=4)
    acc.build_graph(Circuit((ExpGate(MajoranaOperator({(0,): 1.0}, 2)),), initial_state=[0, 1], system_size=2))
    assert acc.n_gates == 1
    acc.build_graph(Circuit((ExpGate(MajoranaOperator({(1,): 1.0}, 2)),), initial_state=[0, 1], system_size=2))
    assert acc.n_gates == 2


def test_circuit_checks_at_propagate():
    p = fixtures.load("rx_rz_ry_rz_exact")
    prop = _prop(p)
    with pytest.raises(TypeError, match="qubit"):
        prop.propagate(Circuit((ExpGate(PauliOperator({"Z": 1.0}, num_qubits=p.n_modes)),), initial_state=(),
                               system_size=p.n_modes))
    gate = ExpGate(MajoranaOperator({(0, 1): 1.0j}, p.n_modes))
    with pytest.raises(ValueError, match="initial_state"):
        prop.propagate(Circuit((gate,), initial_state=(0, 1), system_size=p.n_modes, parameters=(0.1,)))
    prop.propagate(Circuit((gate,), system_size=p.n_modes, parameters=(0.1,)))  # empty: the propagator's state
    herm = MajoranaPropagator(MajoranaOperator({(0, 1): 1.0j}, 2), [0, 1], cutoff=4)
    with pytest.raises(ValueError, match="not Hermitian"):
        herm.propagate(Circuit((ExpGate(MajoranaOperator({(0, 1): 1.0}, 2)),), initial_state=[0, 1], system_size=2,
                               parameters=(0.3,)))


def test_seed_length_and_failed_build_keep_the_axis():
    _small(lower_atol=1e-12).build_graph(_two_gates(), seed_parameters=np.array([0.5, 0.3]))
    prop = _small(lower_atol=1e-12)
    with pytest.raises(RuntimeError, match="length of parameters"):
        prop.build_graph(_two_gates(), seed_parameters=[0.5])
    assert prop.n_parameters == 0
    prop.build_graph(_two_gates(), seed_parameters=[0.5, 0.3])
    assert prop.n_parameters == 2 and min(prop.parameter_mapping) == 0
    ref = _small(lower_atol=1e-12)
    ref.build_graph(_two_gates(), seed_parameters=[0.5, 0.3])
    np.testing.assert_allclose(prop.expval([0.5, 0.3]), ref.expval([0.5, 0.3]))


def test_inplace_contraction_resets_the_axis():
    c2 = Circuit((ExpGate(MajoranaOperator({(0, 1): -1.0j}, 8)),), parameters=(0.4,), system_size=8)
    prop = _small(lower_atol=1e-15)
    prop.build_graph(_two_gates())
    assert prop.n_parameters == 2
    prop.contract_partially([0.5, 0.3])
    assert prop.graph_layers == 0 and prop.n_parameters == 0
    prop.build_graph(c2)
    assert prop.n_parameters == 1 and prop.parameter_mapping == [0]
    ref = _small(lower_atol=1e-15)
    ref.build_graph(_two_gates())
    ref.contract_partially([0.5, 0.3])
    ref.build_graph(c2)
    np.testing.assert_allclose(prop.expval([0.4]), ref.expval([0.4]))


def test_propagate_after_build_graph_rejected():
    c1 = Circuit((ExpGate(MajoranaOperator({(4, 5): -1.0j}, 8)),), initial_state=(), system_size=8, parameters=(0.3,))
    c2 = Circuit((ExpGate(MajoranaOperator({(2, 3): -1.0j}, 8)),), initial_state=(), system_size=8, parameters=(0.4,))
    prop = _small()
    prop.build_graph(c1)
    with pytest.raises(RuntimeError, match="non-empty graph"):
        prop.propagate(c2)


@pytest.mark.parametrize("n_params", [1, 3])
def test_propagate_parameter_length_is_the_engines_check(n_params):
    """A parameter vector shorter or longer than max(parameter_mapping) + 1 reaches the engine's check
    and raises the reference's RuntimeError (Validation.cpp:79-91), for propagate as for build_graph;
    the store is left as it was."""
    from monoprop_amd._engine import DeviceSimulator
    from monoprop_amd.circuit import expand_monomials

    print(f"case: propagate with {n_params} parameters for 2", flush=True)
    c = _two_gates()
    monos, gcs, pm, _ = expand_monomials(c.gates, c.resolved_mapping, 8)
    d = DeviceSimulator(_OBS.terms, 16, [], logical_num_modes=8, basis="majorana")
    with pytest.raises(RuntimeError, match=r"length of parameters \(%d\) must be the same as max\(parameter_mapping\)\+1 \(2\)" % n_params):
        d.propagate(monos, pm, gcs, [0.5, 0.3, 0.1][:n_params])
    assert d.size() == len(_OBS.terms)
    d.propagate(monos, pm, gcs, [0.5, 0.3])
    assert d.size() >= len(_OBS.terms)
