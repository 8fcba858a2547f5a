"""update_initial_operator on the HIP engine vs the oracle, and the reference's own cases through the
public API (tests/test_monoprop_trivial.py:72-134, tests/test_pauli.py:92-99,
tests/test_parameter_validation.py:175-200)."""

from __future__ import annotations

import numpy as np
import pytest

import fixtures
from monoprop_amd import MajoranaPropagator, PauliPropagator, workloads
from monoprop_amd._engine import DeviceSimulator
from monoprop_amd.circuit import Circuit, ExpGate, expand_monomials
from monoprop_amd.majorana import MajoranaOperator
from monoprop_amd.pauli import PauliOperator

pytestmark = pytest.mark.gpu


def _terms_of_rows(o, idx):
    """Slot tuples of store rows `idx` (from the oracle's export, which lists every row at atol 0)."""
    ev = o.evolved_operator(0.0)
    keys = [k for k in ev if k]
    return [keys[i] for i in idx]


def _same(o, d):
    wo, co = o.rows()
    wd, cd = d.rows()
    assert np.array_equal(wo, wd) and np.array_equal(co, cd)


def test_heisenberg_after_propagate(oracle_mod):
    print("case: update after propagate", flush=True)
    n = 20
    circ = workloads.kicked_ising_chain(n, 3)
    monos, gcs, pm, _ = expand_monomials(circ.gates, circ.resolved_mapping, n)
    terms = workloads.z_observable(n, [10]).get_local_operator().terms
    o = oracle_mod.Oracle(terms, n, basis="pauli", cutoff=8, lower_atol=1e-8)
    d = DeviceSimulator(terms, 8, [], lower_atol=1e-8, cutoff_type="support", logical_num_modes=n, basis="pauli")
    o.propagate(monos, pm, gcs, circ.parameters)
    d.propagate(monos, pm, gcs, circ.parameters)
    keys = _terms_of_rows(o, [0, 3, 17, 50])
    new = {k: 0.1 * (i + 1) for i, k in enumerate(keys)}
    new[()] = -0.5
    o.update_initial_operator(new)
    d.update_initial_operator(new)
    _same(o, d)
    assert abs(o.expectation() - d.expectation_value([])) <= 1e-15
    with pytest.raises(RuntimeError, match="not found in the operator"):
        d.update_initial_operator({(0,): 1.0})  # X on qubit 0: outside the light cone of Z_10 after 3 steps
    _same(o, d)


@pytest.mark.parametrize("name", ["lih_fermionic_spin_exact", "rx_rz_ry_rz_exact"])
def test_heisenberg_graph_reweight(oracle_mod, name):
    print(f"case: update graph {name}", flush=True)
    p = fixtures.load(name)
    c = p.circuit()
    monos, gcs, pm, gidx = expand_monomials(c.gates, c.resolved_mapping, p.n_modes)
    o = oracle_mod.Oracle(p.operator.terms, p.n_modes, basis="majorana", cutoff=6, cutoff_type="length",
                          lower_atol=1e-8, initial_state=p.initial_state)
    d = DeviceSimulator(p.operator.terms, 6, p.initial_state, lower_atol=1e-8, cutoff_type="length",
                        logical_num_modes=p.n_modes, basis="majorana")
    o.build_graph(monos, pm, gcs, p.parameters)
    d.build_graph(monos, pm, gcs, gidx, list(p.parameters))
    stale = d.expectation_value_functional()
    keys = _terms_of_rows(o, [0, 1, 5])
    # Hermitian coefficients: w i^{C(|k|,2)}
    new = {k: w * (1j) ** (len(k) * (len(k) - 1) // 2) for k, w in zip(keys, (0.5, -0.25, 1.0))}
    o.update_initial_operator(new)
    d.update_initial_operator(new)
    with pytest.raises(RuntimeError, match="MP object has been modified"):
        stale(list(p.parameters))
    x = np.random.default_rng(1).random(len(p.parameters))
    assert np.array_equal(np.asarray(d.contract_partially(list(x), False)), o.graph_contract(x))
    assert d.expectation_value(list(x)) == o.graph_ev(x)
    vo, go = o.graph_grad(x)
    vd, gd = d.expectation_value_and_gradient(list(x))
    assert vo == vd
    assert np.allclose(gd, go, rtol=1e-10, atol=1e-12)


def test_schrodinger_observable_swap(oracle_mod):
    print("case: update schrodinger", flush=True)
    p = fixtures.load("lih_fermionic_spin_exact")
    c = p.circuit()
    monos, gcs, pm, _ = expand_monomials(c.gates, c.resolved_mapping, p.n_modes)
    kw = dict(basis="majorana", cutoff=6, cutoff_type="length", lower_atol=1e-8, initial_state=p.initial_state,
              schrodinger_cutoff=6)
    o = oracle_mod.Oracle(p.operator.terms, p.n_modes, **kw)
    d = DeviceSimulator(p.operator.terms, 6, p.initial_state, lower_atol=1e-8, cutoff_type="length",
                        logical_num_modes=p.n_modes, basis="majorana", schrodinger_cutoff=6)
    o.propagate(monos, pm, gcs, p.parameters)
    d.propagate(monos, pm, gcs, p.parameters)
    new = {(0, 1): 0.5j, (2, 3, 4, 5): 0.25, (1, 6): 2.0j, (): -1.0}
    o.update_initial_operator(new)
    d.update_initial_operator(new)
    _same(o, d)
    assert o.expectation() == d.expectation_value([])


def test_reference_cases_public_api():
    print("case: update public api", flush=True)
    mp = MajoranaPropagator(MajoranaOperator({(0, 1, 2, 4): 1}, 8), [], cutoff=16)
    assert np.array_equal(mp.contract_partially(inplace=False), [-1.0])
    mp.update_initial_operator(MajoranaOperator({(0, 1, 2, 4): 2.0 + 0j}, 8))
    assert np.array_equal(mp.contract_partially(inplace=False), [-2.0])
    mp = MajoranaPropagator(MajoranaOperator({(0, 3): 1.0j}, 4), [], cutoff=8, schrodinger_cutoff=8)
    assert mp.expval() == 0.0
    mp.update_initial_operator(MajoranaOperator({(0, 1): 2.0j}, 4))
    assert np.isclose(mp.expval(), -2.0)
    mp = PauliPropagator(PauliOperator({"Z": 1.0}, num_qubits=1), initial_state=[], cutoff=1)
    mp.update_initial_operator(PauliOperator({"Z": 2.75}, num_qubits=1))
    assert mp.evolved_operator() == PauliOperator({"Z": 2.75}, num_qubits=1)
    # a functional created before the re-weight refuses to answer; a new one answers for the new operator
    mp = PauliPropagator(PauliOperator({"ZZ": 1.0}, num_qubits=2), [0, 1], cutoff=2)
    mp.build_graph(Circuit((ExpGate(PauliOperator({"XY": 1.0}, 2)), ExpGate(PauliOperator({"YX": 1.0}, 2))), 2))
    fn = mp.expval_functional()
    fn([0.3, 0.7])
    mp.update_initial_operator(PauliOperator({"ZZ": 0.5}, num_qubits=2))
    with pytest.raises(RuntimeError, match="MP object has been modified"):
        fn([0.3, 0.7])
    assert mp.expval_functional()([0.3, 0.7]) == pytest.approx(mp.expval([0.3, 0.7]))
