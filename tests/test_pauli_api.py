"""Pauli / PauliOperator / Pauli circuits on the host, with the cases of the reference's
tests/test_pauli.py:100-400: canonical terms, validation, equality and closeness, pairwise commutation
and the Majorana (slot) view."""

from __future__ import annotations

import pytest

from monoprop_amd import Circuit, ExpGate, MajoranaOperator, Pauli, PauliOperator

P = PauliOperator


def test_pauli_terms():
    p = Pauli("XYZ")
    assert p.string == "XYZ" and p.qubits == (0, 1, 2)
    assert repr(Pauli("IX")) == "Pauli('X', (1,))"
    assert Pauli("IZ", (0, 1)) == Pauli("Z", 1) == Pauli("Z", (1,))
    assert Pauli("XY", (1, 0)) == Pauli("YX", (0, 1))
    assert (Pauli("XY") == Pauli("XY")) is True and (Pauli("XY") == Pauli("XZ")) is False
    for other in ("X", 1, object()):
        assert (Pauli("X", 0) == other) is False
    assert {Pauli("XY", (0, 1)), Pauli("YX", (1, 0))} == {Pauli("XY", (0, 1))}
    with pytest.raises(ValueError, match="Invalid characters"):
        Pauli("XA")
    with pytest.raises(ValueError, match="same length"):
        Pauli("XY", (0,))
    with pytest.raises(ValueError, match="Duplicate qubit indices"):
        Pauli("XY", (0, 0))


def test_operator_construction_and_validation():
    op = P({"XY": 1.0, "IZ": 0.5}, num_qubits=2)
    assert set(op.terms) == {Pauli("XY", (0, 1)), Pauli("Z", 1)} and list(op.terms.values()) == [1.0, 0.5]
    assert op.num_qubits == 2 and P({"XYZ": 1.0}, num_qubits=3).num_qubits == 3
    with pytest.raises(ValueError, match="qubit index"):
        P({"XYZ": 1.0}, num_qubits=2)
    assert len(P({"X": 1.0, "Y": 2.0, "Z": 3.0}, num_qubits=1)) == 3
    single = P({"XIYZ": 1}, num_qubits=4)
    assert single.num_qubits == 4 and len(single) == 1
    with pytest.raises(ValueError, match="Invalid characters"):
        P({"XA": 1.0}, num_qubits=2)
    assert set(P({"XYIZ": 1.0}, num_qubits=4).terms) == {Pauli("XYIZ")}
    with pytest.raises(TypeError, match="num_qubits must be an integer"):
        P._from_terms(["X"], [1.0], num_qubits=None)
    with pytest.raises(TypeError):
        P({"IXYZ": 2.0})
    with pytest.raises(ValueError, match="Operator has complex terms"):
        P({Pauli("X", 0): 1.0j}, num_qubits=1)


def test_operator_str():
    r = str(P({"XY": 1.0}, num_qubits=2))
    assert "PauliOperator" in r and "XY" in r
    r = str(P({Pauli("X", i): 1.0 for i in range(10)}, num_qubits=10))
    assert "PauliOperator" in r and "10 terms" in r


def test_majorana_view():
    ident = P({"I": 2.5}, num_qubits=1).get_majorana_operator()
    assert ident.num_modes == 1 and ident.terms == {(): pytest.approx(2.5)}
    z = P({"Z": 1.0}, num_qubits=1).get_majorana_operator()
    assert z.num_modes == 1 and z.terms == {(0, 1): pytest.approx(-1.0j)}


@pytest.mark.parametrize(("terms", "expected"), [
    ({}, True), ({Pauli("X", 0): 1.0}, True), ({Pauli("XX", (0, 1)): 1.0, Pauli("ZZ", (0, 1)): 1.0}, True),
    ({Pauli("X", 0): 1.0, Pauli("Z", 0): 1.0}, False),
])
def test_all_pairwise_commute(terms, expected):
    assert P(terms, num_qubits=2).all_pairwise_commute() is expected


@pytest.mark.parametrize(("left", "right", "eq", "close"), [
    (P({"XY": 1.0}, 2), P({"XY": 1.0}, 2), True, True),
    (P({"XY": 1.0}, 2), P({"XZ": 1.0}, 2), False, False),
    (P({"XY": 1.0}, 2), P({"XYI": 1.0}, 3), False, False),
    (P({"XY": 0.0}, 2), P({}, 2), False, True),
    (P({"XY": 1.0}, 2), P({"XY": 1.0 + 1e-9}, 2), False, True),
    (P({"XY": 1.0}, 2), P({"XY": 1.1}, 2), False, False),
    (P({"XX": 1e-16}, 2), P({}, 2), False, True),
    (P({"XY": 1.0, "IZ": 0.5}, 2), P({"XY": 1.0}, 2), False, False),
])
def test_equality_and_closeness(left, right, eq, close):
    assert (left == right) is eq
    assert left.isclose(right) is close
    with pytest.raises(TypeError):
        left.isclose("not an operator")


def test_pauli_circuits():
    gate = lambda: ExpGate(P({Pauli("X", 0): 1.0}, num_qubits=1))  # noqa: E731
    c = Circuit((gate(), gate()), initial_state=(0,), system_size=1, parameters=(0.5, 0.5))
    assert len(c) == 2 and c.initial_state == (0,)
    assert len(Circuit((), initial_state=(0,), system_size=1)) == 0
    d = Circuit((gate(), gate()), initial_state=(), system_size=1)
    assert list(d.resolved_mapping) == [0, 1] and d.n_parameters == 2
    with pytest.raises(TypeError, match="mix"):
        Circuit((ExpGate(P({Pauli("XY", [0, 1]): 1.0}, num_qubits=2)), ExpGate(MajoranaOperator({(0, 1): 1.0j}, 2))),
                system_size=2)
    with pytest.raises(ValueError, match=r"The provided generator must be composed of commuting terms."):
        ExpGate(P({Pauli("X", 0): 1.0, Pauli("Z", 0): 1.0}, num_qubits=1))
    ExpGate(P({Pauli("XX", (0, 1)): 1.0, Pauli("ZZ", (0, 1)): 1.0}, num_qubits=2))
