"""only_rotate_len_k through the public API, with the reference's own cases
(tests/test_only_rotate_k.py:48-146, 194-275): None is the uncapped default, a capped Schrodinger
rotation leaves the long terms untouched and rotates the short ones as the uncapped run does, and
out-of-range caps raise before anything runs."""

from __future__ import annotations

from contextlib import nullcontext

import numpy as np
import pytest

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
