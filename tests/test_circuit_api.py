"""Circuit / ExpGate behaviour on the host, with the cases of the reference's tests/test_circuit.py:59-425
(the fermionic-generator cases are out of scope): equality, validation messages, parameter mapping,
dense arrays and composition."""

from __future__ import annotations

import numpy as np
import pytest

import fixtures
from monoprop_amd import Circuit, ExpGate, Majorana, MajoranaOperator, Pauli, PauliOperator
from monoprop_amd.circuit import validate_parameter_mapping


def _m(terms, n=2):
    return MajoranaOperator(terms, num_modes=n)


def test_exp_rejects_bare_terms():
    with pytest.raises(TypeError, match="not a bare term"):
        ExpGate(Majorana(0, 1))
    with pytest.raises(TypeError, match="not a bare term"):
        ExpGate(Pauli("X", 0))


def test_exp_equality_and_repr():
    gen = _m({(0, 1): 1.0j})
    assert ExpGate(gen, index=0) == ExpGate(gen, index=0)
    assert ExpGate(gen, index=0) != ExpGate(gen, index=1)
    assert ExpGate(gen) != "not an ExpGate"
    assert repr(ExpGate(gen, index=0)).startswith("ExpGate(")


@pytest.mark.parametrize(("gen", "size"), [(_m({(0, 1): 1.0j}, 3), 3),
                                           (PauliOperator({Pauli("XX"): 1.0}, num_qubits=2), 2)])
def test_exp_gate_system_size(gen, size):
    assert ExpGate(gen).system_size == size


def test_circuit_equality_and_len():
    gen = _m({(0, 1): 1.0j})
    a = Circuit((ExpGate(gen),), initial_state=(0,), system_size=2, parameters=(0.3,))
    assert a == Circuit((ExpGate(gen),), initial_state=(0,), system_size=2, parameters=(0.3,))
    assert a != Circuit((ExpGate(gen),), initial_state=(0,), system_size=2, parameters=(0.9,))
    assert a != "not a circuit"
    assert len(Circuit((ExpGate(gen), ExpGate(gen)), initial_state=(), system_size=2)) == 2


@pytest.mark.parametrize(("kwargs", "exc", "match"), [
    (dict(gates=("not a gate",), initial_state=(), system_size=0), TypeError, "Circuit gates must be ExpGate"),
    (dict(gates=(), initial_state=(0, 1, 0), system_size=2), ValueError, "Duplicate indices"),
    (dict(gates=(), initial_state=(0, 2), system_size=2), ValueError, "initial_state entries must be in"),
    (dict(gates=(ExpGate(_m({(0, 1): 1.0j})),), initial_state=(), system_size=4), ValueError,
     "does not match circuit system_size"),
    (dict(gates=(ExpGate(_m({(0, 1): 1.0j})), ExpGate(PauliOperator({Pauli("X", 0): 1.0}, num_qubits=2))),
          initial_state=(), system_size=2), TypeError, "cannot mix qubit and Majorana"),
    (dict(gates=(ExpGate(_m({(0, 1): 1.0j}), index=0), ExpGate(_m({(2, 3): 1.0j}), index=2)), initial_state=(),
          system_size=2), ValueError, "contiguous"),
    (dict(gates=(ExpGate(_m({(0, 1): 1.0j}), index=0), ExpGate(_m({(2, 3): 1.0j}))), initial_state=(),
          system_size=2), ValueError, "every gate must set"),
    (dict(gates=(ExpGate(_m({(0, 1): 1.0j})),), initial_state=(), system_size=2, parameters=(0.1, 0.2)),
     ValueError, "1 parameters"),
    (dict(gates=(ExpGate(_m({})), ExpGate(_m({(0, 1): 1.0j}))), initial_state=(), system_size=2,
          parameters=(0.1,)), ValueError, "2 gates"),
])
def test_circuit_validation(kwargs, exc, match):
    gates = kwargs.pop("gates")
    with pytest.raises(exc, match=match):
        Circuit(gates, **kwargs)


def test_identity_gate_dropped_without_parameters():
    c = Circuit((ExpGate(_m({})), ExpGate(_m({(0, 1): 1.0j}))), initial_state=(), system_size=2)
    assert len(c.gates) == 1 and c.parameters == ()


def test_dense_arrays_round_trip_and_errors():
    p = fixtures.load("lih_fermionic_spin_exact")
    c = p.circuit()
    assert [mono for gate in c for mono in gate.generator.terms] == [tuple(m) for m in p.majoranas]
    assert list(c.resolved_mapping) != []
    np.testing.assert_allclose(c.parameters, p.parameters)
    assert c.initial_state == tuple(int(i) for i in p.initial_state)
    with pytest.raises(ValueError, match="must be non-negative"):
        Circuit.from_dense_arrays(majoranas=[[-1, 0]], gen_coeffs=[1.0], param_inds=[0], system_size=2)
    with pytest.raises(ValueError, match="acts on an index >= 2\\*system_size"):
        Circuit.from_dense_arrays(majoranas=[[0, 4]], gen_coeffs=[1.0], param_inds=[0], system_size=2)
    empty = Circuit.from_dense_arrays(majoranas=[], gen_coeffs=[], param_inds=[], system_size=2)
    assert empty.gates == () and empty.family == "empty"


def test_mappings():
    c = Circuit((ExpGate(_m({(0, 1): 1.0j})), ExpGate(_m({(2, 3): 1.0j}))), initial_state=(), system_size=2)
    assert c.resolved_mapping == (0, 1) and c.n_parameters == 2
    tied = Circuit((ExpGate(_m({(0, 1): 1.0j}), index=0), ExpGate(_m({(2, 3): 1.0j}), index=1),
                    ExpGate(_m({(0, 3): 1.0j}), index=0)), initial_state=(), system_size=2)
    assert tied.resolved_mapping == (0, 1, 0) and tied.n_parameters == 2
    with pytest.raises(ValueError, match="3 entries but there are 2 gates"):
        validate_parameter_mapping((0, 1, 0), expected_len=2, unit="gates")
    with pytest.raises(ValueError, match="contiguous"):
        validate_parameter_mapping((0, 2), expected_len=2, unit="graph layers")
    validate_parameter_mapping((0, 1, 0), expected_len=3, unit="gates")


def test_composition():
    a = Circuit((ExpGate(_m({(0,): 1.0})), ExpGate(_m({(1,): 1.0}))), initial_state=(), system_size=2,
                parameters=(0.1, 0.2))
    b = Circuit((ExpGate(_m({(2,): 1.0})),), initial_state=(), system_size=2, parameters=(0.3,))
    ab = a + b
    assert ab.resolved_mapping == (0, 1, 2) and ab.parameters == (0.1, 0.2, 0.3) and ab.n_parameters == 3
    q = Circuit((ExpGate(PauliOperator({Pauli("X", 0): 1.0}, num_qubits=2)),), system_size=2)
    with pytest.raises(TypeError, match="gate families differ"):
        _ = Circuit((ExpGate(_m({(0, 1): 1.0j})),), system_size=2) + q
    with pytest.raises(ValueError, match="different system_size"):
        _ = (Circuit((ExpGate(_m({(0,): 1.0})),), system_size=2)
             + Circuit((ExpGate(_m({(0,): 1.0}, 4)),), system_size=4))
    with pytest.raises(ValueError, match="different initial states"):
        _ = (Circuit((ExpGate(_m({(0,): 1.0})),), initial_state=(0,), system_size=2)
             + Circuit((ExpGate(_m({(1,): 1.0})),), initial_state=(1,), system_size=2))
