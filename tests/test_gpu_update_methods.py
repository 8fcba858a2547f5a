"""Truncation setters through the public API with the reference's cases (tests/test_update_methods.py:
27-230; messages of cpp/include/monoprop/MonomialPropagator.h:168-201): valid and edge values, the
ordering rule between lower_atol and upper_atol, type errors as the binding raises them, a rejected
value leaving the old one in place, and evolutions after an update keeping a different term set."""

from __future__ import annotations

import pytest

import fixtures
from monoprop_amd import Circuit, ExpGate, MajoranaOperator, MajoranaPropagator

pytestmark = pytest.mark.gpu


@pytest.fixture
def mp():
    return MajoranaPropagator(initial_operator=MajoranaOperator({(0, 1): 1.0j}, num_modes=4), initial_state=[],
                              cutoff=4, lower_atol=1e-8, upper_atol=1e-4)


@pytest.mark.parametrize("value", [1e-10, None])
def test_lower_atol_valid(mp, value):
    mp.lower_atol = value
    assert mp.lower_atol == value


def test_lower_atol_above_upper_rejected_and_kept(mp):
    mp.lower_atol = 1e-10
    with pytest.raises(RuntimeError, match=r"New lower_atol \(0\.001\)"):
        mp.lower_atol = 1e-3
    assert mp.lower_atol == 1e-10


def test_lower_atol_edges(mp):
    mp.lower_atol = 1e-4  # equal to upper
    assert mp.lower_atol == 1e-4
    mp.lower_atol = None
    mp.upper_atol = 1e-15
    mp.lower_atol = 1e-16
    assert mp.lower_atol == 1e-16


@pytest.mark.parametrize("name", ["lower_atol", "upper_atol"])
@pytest.mark.parametrize("bad", ["invalid", [1e-8]])
def test_atol_type_errors(mp, name, bad):
    with pytest.raises(TypeError):
        setattr(mp, name, bad)


@pytest.mark.parametrize("value", [1e-3, None, 1e-8, 1.0])
def test_upper_atol_valid_and_edges(mp, value):
    mp.upper_atol = value
    assert mp.upper_atol == value


def test_upper_atol_below_lower_rejected_and_kept(mp):
    mp.upper_atol = 1e-3
    with pytest.raises((ValueError, RuntimeError), match="must be greater than or equal to"):
        mp.upper_atol = 1e-9
    assert mp.upper_atol == 1e-3


def test_atol_consistency_and_ordering(mp):
    mp.lower_atol = None
    mp.upper_atol = None
    assert mp.lower_atol is None and mp.upper_atol is None
    mp.upper_atol = 1e-5
    mp.lower_atol = 1e-6
    assert (mp.lower_atol, mp.upper_atol) == (1e-6, 1e-5)
    mp.upper_atol = 1e-3
    mp.lower_atol = 1e-4
    assert (mp.lower_atol, mp.upper_atol) == (1e-4, 1e-3)
    for order in ("lower_first", "upper_first"):
        mp.lower_atol = None
        mp.upper_atol = None
        if order == "lower_first":
            mp.lower_atol = 1e-10
            mp.upper_atol = 1e-6
        else:
            mp.upper_atol = 1e-6
            mp.lower_atol = 1e-10
        assert (mp.lower_atol, mp.upper_atol) == (1e-10, 1e-6)


def test_cutoff_values_and_types(mp):
    mp.cutoff = 6
    gate = ExpGate(MajoranaOperator({(0, 1, 2, 3, 4, 5): 1.0j}, num_modes=4))
    mp.build_graph(Circuit((gate,), initial_state=(), system_size=4))
    assert mp.size() > 0
    for edge in (0, 1000):
        mp.cutoff = edge
        assert mp.cutoff == edge
    with pytest.raises((ValueError, TypeError)):
        mp.cutoff = -1
    for bad in (3.14, "invalid"):
        with pytest.raises(TypeError):
            mp.cutoff = bad
    assert mp.cutoff == 1000


def test_cutoff_type_values_and_types(mp):
    for t in ("support", "length"):
        mp.cutoff_type = t
        assert mp.cutoff_type == t
    with pytest.raises((ValueError, RuntimeError)):
        mp.cutoff_type = "invalid_type"
    for bad in (None, 123):
        with pytest.raises((ValueError, RuntimeError, TypeError)):
            mp.cutoff_type = bad
    assert mp.cutoff_type == "length"


def test_integration_and_reset(mp):
    seq = Circuit.from_dense_arrays(initial_state=[], majoranas=[(0, 2), (1, 3)], gen_coeffs=[0.0, 0.0],
                                    param_inds=[0, 1], system_size=4, parameters=[1.0, 1.0])
    p = MajoranaPropagator(initial_operator=MajoranaOperator({(0, 1): 1.0j}, num_modes=4),
                           initial_state=seq.initial_state, cutoff=4)
    for attr, value in {"lower_atol": 1e-10, "upper_atol": 1e-6, "cutoff": 6, "cutoff_type": "support"}.items():
        setattr(p, attr, value)
    p.propagate(seq)
    assert isinstance(p.expval(), (int, float))
    for attr, value in {"lower_atol": 1e-12, "upper_atol": 1e-2, "cutoff": 10, "cutoff_type": "length"}.items():
        setattr(mp, attr, value)
    for attr, value in {"lower_atol": None, "upper_atol": None, "cutoff_type": "support", "cutoff": 4}.items():
        setattr(mp, attr, value)
    assert (mp.lower_atol, mp.upper_atol, mp.cutoff_type, mp.cutoff) == (None, None, "support", 4)


@pytest.mark.parametrize("test_type", ["cutoff", "lower_atol", "upper_atol"])
def test_evolutions_after_updates(test_type):
    print(f"case: update {test_type}", flush=True)
    prob = fixtures.load("lih_fermionic_spin_exact")
    circuit = prob.circuit()
    cutoff = 2 if test_type == "upper_atol" else 6
    a = MajoranaPropagator(prob.operator, prob.initial_state, cutoff=cutoff)
    a.propagate(circuit)
    b = MajoranaPropagator(prob.operator, prob.initial_state, cutoff=cutoff)
    if test_type == "lower_atol":
        b.lower_atol = 1
    elif test_type == "upper_atol":
        b.upper_atol = 0.001
    else:
        b.cutoff = 4
    b.propagate(circuit)
    assert b.size() != a.size()
