"""Majorana / MajoranaOperator behaviour on the host, with the cases of the reference's
tests/test_majorana.py:24-230: term canonicalisation and its sign, validation, equality, closeness,
pairwise commutation and the dense-array grouping."""

from __future__ import annotations

import pytest

from monoprop_amd import Circuit, Majorana, MajoranaOperator


def test_majorana_terms():
    assert Majorana(1, 4, 5).indices == (1, 4, 5)
    with pytest.raises(ValueError, match="sorted"):
        Majorana(5, 4, 1)
    assert Majorana().indices == ()
    with pytest.raises(ValueError, match="non-negative"):
        Majorana(0, -1)
    with pytest.raises(ValueError, match="distinct"):
        Majorana(2, 2)
    assert Majorana.from_unsorted(5, 4, 1) == (Majorana(1, 4, 5), -1.0)
    assert Majorana.from_unsorted(2, 0, 1) == (Majorana(0, 1, 2), 1.0)
    assert Majorana.from_unsorted(3, 1, 1, 3) == (Majorana(), 1.0)
    assert Majorana.from_unsorted(1, 2, 1) == (Majorana(2), -1.0)
    a, b = Majorana(4, 5), Majorana(4, 5)
    assert a == b and hash(a) == hash(b) and {a, b} == {a}
    assert (Majorana(0, 1) == (0, 1)) is False and Majorana(0, 1) != (0, 1)
    assert repr(Majorana(4, 5)) == "Majorana(4, 5)"


def test_operator_keys_validated_and_normalised():
    assert MajoranaOperator({(0, 0): 1.0}, num_modes=2).terms == {(): 1.0}
    with pytest.raises(ValueError, match="non-negative"):
        MajoranaOperator({(-1, 0): 1.0}, num_modes=2)
    with pytest.raises(ValueError, match="acts on an index >= num_modes=2"):
        MajoranaOperator({(0, 4): 1.0}, num_modes=2)
    assert MajoranaOperator({(0, 3): 1.0}, num_modes=2).terms == {(0, 3): 1.0}
    op = MajoranaOperator({(1, 0): 1.0, (0, 1): 2.0, (2, 3): 1e-15}, num_modes=4)
    assert op.terms == {(0, 1): 1.0, (2, 3): 1e-15} and len(op) == 2 and op.num_modes == 4
    rt = MajoranaOperator({(0, 1): 1.0j, (2, 3): -0.5}, num_modes=4)
    assert rt.terms == {(0, 1): 1.0j, (2, 3): -0.5} and rt.get_majorana_operator() is rt


def test_dense_arrays_group_by_parameter():
    c = Circuit.from_dense_arrays(majoranas=[(0, 1), (2, 3), (0, 3)], gen_coeffs=[0.5, -0.5, 1.0],
                                  param_inds=[0, 0, 1], system_size=2, parameters=[1.0, 2.0], initial_state=[0, 1])
    assert c.initial_state == (0, 1) and c.parameters == (1.0, 2.0) and c.resolved_mapping == (0, 1)
    assert len(c) == 2 and len(c.gates[0].generator.terms) == 2 and len(c.gates[1].generator.terms) == 1


@pytest.mark.parametrize(("terms", "expected"), [
    ({}, True), ({(0,): 1.0}, True), ({(0, 1): 1.0j, (2, 3): 1.0j}, True),
    ({(0, 2, 3, 5): 1.0, (0, 2, 4, 6): 1.0}, True), ({(0,): 1.0, (1,): 1.0}, False),
    ({(0,): 1.0, (0, 1): 1.0}, False),
])
def test_all_pairwise_commute(terms, expected):
    assert MajoranaOperator(terms, num_modes=10).all_pairwise_commute() is expected


def test_equality_and_closeness():
    m = MajoranaOperator
    assert m({(0, 1): 1.0j}, 2) == m({(0, 1): 1.0j}, 2)
    assert (m({(0, 1): 1.0j}, 2) == m({(0, 1): 1.0}, 2)) is False
    assert (m({(0, 1): 1.0}, 2) == m({(0, 1): 1.0}, 3)) is False
    assert (m({(0, 1): 0.0}, 2) == m({}, 2)) is False
    assert m({(0, 1): 1.0}, 2).isclose(m({(0, 1): 1.0 + 1e-9}, 2)) is True
    assert m({(0, 1): 1.0}, 2).isclose(m({(0, 1): 1.1}, 2)) is False
    assert m({(0, 1): 1.0, (2, 3): 1e-12}, 4).isclose(m({(0, 1): 1.0}, 4)) is True
    assert m({(0, 1): 1.0}, 2).isclose(m({(0, 1): 1.0}, 3)) is False
    with pytest.raises(TypeError, match="Cannot compare MajoranaOperator with dict"):
        m({(0, 1): 1.0}, 2).isclose({(0, 1): 1.0})


@pytest.mark.parametrize(("monos", "coeffs", "n"), [
    ([(0, 1)], [1.0], 4), ([(i, i + 1) for i in range(0, 16, 2)], [1.0] * 8, 8),
    ([(i, i + 1) for i in range(0, 18, 2)], [float(i) for i in range(0, 18, 2)], 9),
])
def test_str(monos, coeffs, n):
    op = MajoranaOperator(dict(zip(monos, coeffs)), n)
    r = str(op)
    assert r.startswith("MajoranaOperator(") and f"{len(op.terms)} terms" in r
