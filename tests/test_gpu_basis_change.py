"""The cutoff basis (basis_change; binder.h:71,185-189, MonomialPropagator.h:204-211,
MonomialPropagator.inl:461-495) on the HIP engine: bit-exact against the oracle in the fused propagate
pipeline and in the per-gate graph build, for the Jordan-Wigner table and for random tables; and the
reference's own cases through the public API (tests/test_basis.py:44-89,
cpp/tests/ctor_validation_tests.cpp:153-169)."""

from __future__ import annotations

import copy

import numpy as np
import pytest

from monoprop_amd import (Circuit, MajoranaOperator, MajoranaPropagator, PauliOperator, PauliPropagator,
                          jordan_wigner_basis_change, workloads)
from monoprop_amd._engine import DeviceSimulator
from monoprop_amd.circuit import expand_monomials

pytestmark = pytest.mark.gpu


def _random_table(n_modes: int, seed: int):
    rng = np.random.default_rng(seed)
    return [sorted(int(i) for i in rng.choice(2 * n_modes, size=int(rng.integers(1, 5)), replace=False))
            for _ in range(2 * n_modes)]


def _pair(oracle_mod, p_terms, n_modes, state, table, cutoff, cutoff_type, lower):
    o = oracle_mod.Oracle(p_terms, n_modes, basis="majorana", cutoff=cutoff, cutoff_type=cutoff_type,
                          lower_atol=lower, initial_state=state, basis_change=table)
    d = DeviceSimulator(p_terms, cutoff, state, lower_atol=lower, cutoff_type=cutoff_type,
                        logical_num_modes=n_modes, basis="majorana", basis_change=table)
    return o, d


@pytest.mark.parametrize("table_kind,cutoff_type,cutoff", [("jw", "support", 4), ("jw", "length", 6),
                                                           ("random", "support", 4), ("random", "length", 5)])
def test_propagate_and_graph_match_oracle(oracle_mod, table_kind, cutoff_type, cutoff):
    print(f"case: basis change {table_kind} {cutoff_type} {cutoff}", flush=True)
    circ, n_modes = workloads.hubbard_lattice(2, 2)
    table = jordan_wigner_basis_change(n_modes) if table_kind == "jw" else _random_table(n_modes, 7)
    obs = MajoranaOperator({(0, 1): 0.5j, (2, 3, 4, 5): 0.25, (): 0.5}, n_modes).terms
    state = [0, 3, 4, 7]
    monos, gcs, pm, gidx = expand_monomials(circ.gates, circ.resolved_mapping, n_modes)
    o, d = _pair(oracle_mod, obs, n_modes, state, table, cutoff, cutoff_type, 1e-9)
    o.propagate(monos, pm, gcs, circ.parameters)
    d.propagate(monos, pm, gcs, circ.parameters)
    wo, co = o.rows()
    wd, cd = d.rows()
    assert np.array_equal(wo, wd) and np.array_equal(co, cd)
    # the stored terms are bit-exact; <O> is a device reduction, equal up to summation order
    assert o.expectation() == d.expectation_value([])
    # the native basis keeps a different set here, so the table is in effect
    n_native = oracle_mod.Oracle(obs, n_modes, basis="majorana", cutoff=cutoff, cutoff_type=cutoff_type,
                                 lower_atol=1e-9, initial_state=state)
    n_native.propagate(monos, pm, gcs, circ.parameters)
    assert n_native.size() != o.size()
    # the per-gate pipeline (graph build) applies the same test
    og, dg = _pair(oracle_mod, obs, n_modes, state, table, cutoff, cutoff_type, 1e-9)
    og.build_graph(monos, pm, gcs, circ.parameters)
    dg.build_graph(monos, pm, gcs, gidx, list(circ.parameters))
    x = np.random.default_rng(3).uniform(-1, 1, len(circ.parameters))
    assert np.array_equal(np.asarray(dg.contract_partially(list(x), False)), og.graph_contract(x))


def test_setter_getter_none_and_deepcopy(oracle_mod):
    print("case: basis change setter", flush=True)
    circ, n_modes = workloads.hubbard_lattice(2, 1)
    monos, gcs, pm, _ = expand_monomials(circ.gates, circ.resolved_mapping, n_modes)
    obs = MajoranaOperator({(0, 1): 0.5j}, n_modes).terms
    jw = jordan_wigner_basis_change(n_modes)
    d = DeviceSimulator(obs, 4, [0, 3], lower_atol=1e-9, cutoff_type="support", logical_num_modes=n_modes,
                        basis="majorana")
    assert d.basis_change is None
    d.basis_change = jw
    assert d.basis_change == [list(r) for r in jw]
    e = copy.deepcopy(d)  # the copy owns its table
    del d
    o = oracle_mod.Oracle(obs, n_modes, basis="majorana", cutoff=4, cutoff_type="support", lower_atol=1e-9,
                          initial_state=[0, 3], basis_change=jw)
    o.propagate(monos, pm, gcs, circ.parameters)
    e.propagate(monos, pm, gcs, circ.parameters)
    assert np.array_equal(o.rows()[0], e.rows()[0]) and np.array_equal(o.rows()[1], e.rows()[1])
    e.basis_change = None  # native again for the gates that follow
    assert e.basis_change is None
    o.set_basis_change(None)
    o.propagate(monos, pm, gcs, circ.parameters)
    e.propagate(monos, pm, gcs, circ.parameters)
    assert np.array_equal(o.rows()[0], e.rows()[0]) and np.array_equal(o.rows()[1], e.rows()[1])


def test_reference_cases_public_api():
    print("case: basis change public api", flush=True)
    n = 6
    mp = MajoranaPropagator(MajoranaOperator({(0,): 1.0}, n), [], cutoff=n // 2)
    mp._simulator.basis_change = jordan_wigner_basis_change(n)
    mp.propagate(Circuit.from_dense_arrays(majoranas=[(5,)], gen_coeffs=[-1.0], param_inds=[0], parameters=[1.0],
                                           system_size=n))
    ev = mp.evolved_operator()
    assert list(ev.terms) == [(0,)]
    assert np.isclose(ev.terms[(0,)].real, np.cos(2 * 1.0))
    for bad, match in (([[0]], "exactly 2\\*logical_num_modes"),
                       ([[0]] * (2 * n - 1), "exactly 2\\*logical_num_modes"),
                       ([[2 * n]] * (2 * n), "out of range")):
        mp2 = MajoranaPropagator(MajoranaOperator({(0,): 1.0}, n), [], cutoff=n // 2)
        with pytest.raises((ValueError, RuntimeError), match=match):
            mp2._simulator.basis_change = bad
        assert mp2._simulator.basis_change is None  # nothing changed
    identity = [[i] for i in range(2 * n)]
    mp3 = MajoranaPropagator(MajoranaOperator({(0,): 1.0}, n), [], cutoff=n // 2)
    mp3._simulator.basis_change = identity
    pp = PauliPropagator(PauliOperator({"ZZ": 1.0}, num_qubits=2), [], cutoff=2)
    with pytest.raises(ValueError, match="does not accept a basis_change"):
        pp._simulator.basis_change = jordan_wigner_basis_change(2)
    with pytest.raises(ValueError, match="requires cutoff_type == Support"):
        pp._simulator.cutoff_type = "length"
