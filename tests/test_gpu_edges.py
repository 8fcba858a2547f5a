"""Edge cases of the propagate path on the HIP engine, against the oracle where there is a result:
the widest rows (256 qubits / 256 modes, 8 words), empty and core-only operators, empty circuits,
identity generators, and the error paths the reference raises on (out-of-range indices, short
parameter vectors, propagate on top of a stored graph)."""

from __future__ import annotations

import numpy as np
import pytest

from monoprop_amd import PauliPropagator
from monoprop_amd._engine import DeviceSimulator
from monoprop_amd.circuit import Circuit, ExpGate, expand_monomials
from monoprop_amd.majorana import MajoranaOperator
from monoprop_amd.pauli import Pauli, PauliOperator

pytestmark = pytest.mark.gpu


def _both(oracle_mod, basis, n, terms, monos, pm, gcs, params, **kw):
    o = oracle_mod.Oracle(terms, n, basis=basis, initial_state=kw.get("state", ()), cutoff=kw["cutoff"],
                          cutoff_type=kw["cutoff_type"], lower_atol=kw.get("lower"),
                          schrodinger_cutoff=kw.get("schr"))
    d = DeviceSimulator(terms, kw["cutoff"], list(kw.get("state", ())), lower_atol=kw.get("lower"),
                        cutoff_type=kw["cutoff_type"], logical_num_modes=n, basis=basis,
                        schrodinger_cutoff=kw.get("schr"))
    o.propagate(monos, pm, gcs, params)
    d.propagate(monos, pm, gcs, params)
    return o, d


def _same(o, d):
    wo, co = o.rows()
    wd, cd = d.rows()
    assert np.array_equal(wo, wd) and np.array_equal(co, cd)
    assert o.expectation() == d.expectation_value([])


@pytest.mark.parametrize("schr", [None, 2])
def test_widest_pauli_rows(oracle_mod, schr):
    """256 qubits: 4 limbs per plane, 8 words per row; generators across every limb boundary."""
    print(f"case: pauli n=256 schr={schr}", flush=True)
    n = 256
    rng = np.random.default_rng(256)
    gates, params = [], []
    for _ in range(60):
        w = int(rng.integers(1, 4))
        qs = sorted(int(q) for q in rng.choice(n, w, replace=False))
        gates.append(ExpGate(PauliOperator({Pauli("".join(rng.choice(list("XYZ"), w)), qs): 1.0}, n)))
        params.append(float(rng.uniform(-np.pi, np.pi)))
    for q in (63, 127, 191):  # bonds across limb boundaries
        gates.append(ExpGate(PauliOperator({Pauli("XY", (q, q + 1)): 1.0}, n)))
        params.append(0.37)
    circ = Circuit(gates, n, params)
    monos, gcs, pm, _ = expand_monomials(circ.gates, circ.resolved_mapping, n)
    obs = PauliOperator({Pauli("Z", (0,)): 0.5, Pauli("ZZ", (64, 200)): -0.25, Pauli("X", (255,)): 1.0}, n)
    o, d = _both(oracle_mod, "pauli", n, obs.get_local_operator().terms, monos, pm, gcs, circ.parameters,
                 cutoff=4, cutoff_type="support", lower=1e-6, state=[1, 64, 130, 255], schr=None if schr is None
                 else 2 * schr)
    assert d.rows()[0].shape[1] == 8
    _same(o, d)


def test_widest_majorana_rows(oracle_mod):
    """256 modes: 512 slots, 8 words per row, odd generators exercise the parity column."""
    print("case: majorana n=256", flush=True)
    n = 256
    rng = np.random.default_rng(512)
    gates, params = [], []
    for _ in range(50):
        w = int(rng.choice([2, 3, 4]))
        idx = tuple(sorted(int(i) for i in rng.choice(2 * n, w, replace=False)))
        gates.append(ExpGate(MajoranaOperator({idx: {2: 1j, 3: 1j, 4: 1.0}[w]}, n)))
        params.append(float(rng.uniform(-np.pi, np.pi)))
    circ = Circuit(gates, n, params)
    monos, gcs, pm, _ = expand_monomials(circ.gates, circ.resolved_mapping, n)
    obs = MajoranaOperator({(0, 1): 0.5j, (300, 301, 508, 511): 0.25, (): 0.5}, n)
    o, d = _both(oracle_mod, "majorana", n, obs.terms, monos, pm, gcs, circ.parameters, cutoff=4,
                 cutoff_type="length", lower=1e-5, state=list(range(0, n, 3)))
    assert d.rows()[0].shape[1] == 8
    _same(o, d)


def test_empty_and_core_only_operators(oracle_mod):
    print("case: empty / core-only operators", flush=True)
    n = 6
    circ = Circuit([ExpGate(PauliOperator({Pauli("XY", (0, 1)): 1.0}, n)),
                    ExpGate(PauliOperator({Pauli("Z", (3,)): 1.0}, n))], n, [0.3, 0.2])
    monos, gcs, pm, _ = expand_monomials(circ.gates, circ.resolved_mapping, n)
    for terms in ({}, {(): 0.75}):
        o, d = _both(oracle_mod, "pauli", n, terms, monos, pm, gcs, circ.parameters, cutoff=4,
                     cutoff_type="support")
        assert d.size() == 0
        assert d.expectation_value([]) == o.expectation() == (0.75 if terms else 0.0)
        assert d.evolved_operator([], 1e-12) == ({(): 0.75} if terms else {})


def test_empty_circuit_and_identity_generator(oracle_mod):
    print("case: empty circuit / identity generator", flush=True)
    n = 4
    terms = {(0,): 1.0, (2, 3): 0.5}
    d = DeviceSimulator(terms, 8, [], logical_num_modes=n, basis="pauli", cutoff_type="support")
    d.propagate([], [], [], [])
    assert d.size() == 2
    # an empty generator monomial rotates nothing (Scan.h:272-274)
    o, d = _both(oracle_mod, "pauli", n, terms, [(), (0, 1)], [0, 1], [1.0, 1.0], [0.4, 0.9], cutoff=8,
                 cutoff_type="support")
    _same(o, d)


def test_error_paths():
    print("case: error paths", flush=True)
    n = 4
    with pytest.raises(Exception, match="out of range"):
        DeviceSimulator({(9,): 1.0}, 4, [], logical_num_modes=n, basis="pauli", cutoff_type="support")
    d = DeviceSimulator({(0,): 1.0}, 4, [], logical_num_modes=n, basis="pauli", cutoff_type="support")
    with pytest.raises(Exception):
        d.propagate([(0, 9)], [0], [1.0], [0.1])  # generator slot past 2n
    with pytest.raises(RuntimeError, match=r"The length of parameters \(1\) must be the same as "
                                           r"max\(parameter_mapping\)\+1 \(4\)\."):
        d.propagate([(2,)], [3], [1.0], [0.1])  # shorter than max(mapping) + 1 (Validation.cpp:79-91)
    with pytest.raises(RuntimeError, match=r"The length of parameters \(3\) must be the same as "
                                           r"max\(parameter_mapping\)\+1 \(1\)\."):
        d.propagate([(2,)], [0], [1.0], [0.1, 0.2, 0.3])  # longer
    d.build_graph([(2,)], [0], [1.0])
    with pytest.raises(RuntimeError, match="graph"):
        d.propagate([(2,)], [0], [1.0], [0.1])
    with pytest.raises(ValueError):
        DeviceSimulator({(0,): 1.0}, 4, [], logical_num_modes=n, basis="pauli", schrodinger_cutoff=-1)
    with pytest.raises(ValueError):
        PauliPropagator(PauliOperator({"Z": 1.0}, 300), [], cutoff=4)  # past 256 qubits


@pytest.mark.parametrize("launch", [(640, 4), (2560, 7), (64, 2)])
def test_launch_parameters_do_not_change_results(oracle_mod, launch):
    """The autotuner's knobs (resolve grid, smallest count group) leave the store bit-exact."""
    print(f"case: launch {launch}", flush=True)
    from monoprop_amd import workloads

    circ = workloads.kicked_ising_lattice(6, 6, theta_h=0.4)
    monos, gcs, pm, _ = expand_monomials(circ.gates, circ.resolved_mapping, 36)
    terms = workloads.mean_z(36).get_local_operator().terms
    o = oracle_mod.Oracle(terms, 36, basis="pauli", cutoff=8, cutoff_type="support", lower_atol=1e-6)
    o.propagate(monos, pm, gcs, circ.parameters)
    d = DeviceSimulator(terms, 8, [], lower_atol=1e-6, cutoff_type="support", logical_num_modes=36, basis="pauli")
    d.set_launch(*launch)
    d.propagate(monos, pm, gcs, circ.parameters)
    assert o.size() > 50000
    _same(o, d)


def test_memory_probes():
    """operator_memory_breakdown / operator_memory_bytes / graph_memory_bytes (binder.h:247-273)."""
    print("case: memory probes", flush=True)
    n = 20
    from monoprop_amd import workloads

    circ = workloads.kicked_ising_chain(n, 3)
    monos, gcs, pm, gidx = expand_monomials(circ.gates, circ.resolved_mapping, n)
    d = DeviceSimulator(workloads.z_observable(n, [10]).get_local_operator().terms, 8, [], lower_atol=1e-8,
                        cutoff_type="support", logical_num_modes=n, basis="pauli")
    d.propagate(monos, pm, gcs, circ.parameters)
    b = d.operator_memory_breakdown()
    parts = ("operator_terms_bytes", "op_coeffs_bytes", "state_coeffs_bytes", "indexing_bytes",
             "init_operator_bytes", "initial_state_bytes", "inverted_index_bytes")
    assert b["total_bytes"] == sum(b[k] for k in parts) == d.operator_memory_bytes()
    assert b["operator_terms_bytes"] >= d.size() * 16 and b["op_coeffs_bytes"] >= d.size() * 8
    assert d.graph_memory_bytes() == 0
    g = DeviceSimulator(workloads.z_observable(n, [10]).get_local_operator().terms, 8, [], lower_atol=1e-8,
                        cutoff_type="support", logical_num_modes=n, basis="pauli")
    g.build_graph(monos, pm, gcs, gidx)
    assert g.graph_memory_bytes() > 0
