"""HIP engine vs the oracle: bit-exact term store and coefficients, on the same inputs.

The store of both engines is append-only in the reference's row order, so parity is array equality
of the raw rows (uint64 words) and of the FP64 coefficients, plus <O> to the bit (both add its terms in ascending row order).
"""

from __future__ import annotations

import numpy as np
import pytest

import fixtures
from monoprop_amd import workloads
from monoprop_amd._engine import DeviceSimulator
from monoprop_amd.circuit import Circuit, ExpGate, expand_monomials
from monoprop_amd.majorana import MajoranaOperator
from monoprop_amd.pauli import Pauli, PauliOperator

pytestmark = pytest.mark.gpu


def _run_both(oracle_mod, basis, n, init_terms, circ, *, cutoff, cutoff_type, lower=None, upper=None,
              state=(), only_rot=None):
    monos, gcs, pm, _ = expand_monomials(circ.gates, circ.resolved_mapping, n)
    o = oracle_mod.Oracle(init_terms, n, basis=basis, cutoff=cutoff, cutoff_type=cutoff_type, lower_atol=lower,
                          upper_atol=upper, initial_state=state)
    o.propagate(monos, pm, gcs, circ.parameters, only_rotate_len_k=only_rot)
    d = DeviceSimulator(init_terms, cutoff, list(state), lower_atol=lower, upper_atol=upper, cutoff_type=cutoff_type,
                        logical_num_modes=n, basis=basis)
    d.propagate(monos, pm, gcs, circ.parameters, only_rot)
    return o, d


def _assert_same(o, d):
    wo, co = o.rows()
    wd, cd = d.rows()
    assert wo.shape == wd.shape, (wo.shape, wd.shape)
    assert np.array_equal(wo, wd)
    assert np.array_equal(co, cd), np.max(np.abs(co - cd))
    no, ao, eo, io = o.gate_stats()
    nd, ad, ed, idd = d.gate_stats()
    assert np.array_equal(no, nd) and np.array_equal(ao, ad) and np.array_equal(eo, ed) and np.array_equal(io, idd)
    assert o.expectation() == d.expectation_value([])


def test_kicked_ising_chain_c1(oracle_mod):
    """20-qubit chain, 5 steps, cutoff 8, lower_atol 1e-8, observable Z_10."""
    print("case: chain n=20 steps=5", flush=True)
    circ = workloads.kicked_ising_chain(20, 5)
    init = workloads.z_observable(20, [10]).get_local_operator().terms
    o, d = _run_both(oracle_mod, "pauli", 20, init, circ, cutoff=8, cutoff_type="support", lower=1e-8)
    assert o.size() > 10000
    _assert_same(o, d)


@pytest.mark.parametrize("n", [7, 70, 130])
@pytest.mark.parametrize("lower,upper,cutoff", [(None, None, 3), (1e-3, None, 4), (1e-4, 1e-2, 2), (None, 0.2, 3)])
def test_random_pauli_circuits(oracle_mod, n, lower, upper, cutoff):
    """Random 1-3 qubit Pauli generators over 1, 2 and 3 limbs per plane."""
    print(f"case: random pauli n={n} lower={lower} upper={upper} cutoff={cutoff}", flush=True)
    rng = np.random.default_rng(n * 31 + cutoff)
    gates, params = [], []
    for _ in range(60):
        w = int(rng.integers(1, 4))
        qs = sorted(int(q) for q in rng.choice(n, w, replace=False))
        letters = "".join(str(rng.choice(list("XYZ"))) for _ in qs)
        gates.append(ExpGate(PauliOperator({Pauli(letters, qs): float(rng.choice([1.0, -0.5, 2.0]))}, n)))
        params.append(float(rng.uniform(-np.pi, np.pi)))
    gates.append(ExpGate(PauliOperator({Pauli("ZZ", (0, n - 1)): 1.0}, n)))
    params.append(-np.pi / 4)
    circ = Circuit(gates, n, params)
    obs = PauliOperator({Pauli("Z", (n // 2,)): 1.0, Pauli("XY", (0, n - 1)): -0.3}, n)
    state = [q for q in range(n) if q % 3 == 0]
    o, d = _run_both(oracle_mod, "pauli", n, obs.get_local_operator().terms, circ, cutoff=cutoff,
                     cutoff_type="support", lower=lower, upper=upper, state=state)
    _assert_same(o, d)


@pytest.mark.parametrize("name", fixtures.EXACT_CASES)
@pytest.mark.parametrize("cutoff,lower", [(None, None), (4, None), (6, 1e-6)])
def test_majorana_fixtures(oracle_mod, name, cutoff, lower):
    print(f"case: fixture {name} cutoff={cutoff} lower={lower}", flush=True)
    p = fixtures.load(name)
    c = cutoff if cutoff is not None else 2 * p.n_modes
    o, d = _run_both(oracle_mod, "majorana", p.n_modes, p.operator.terms, p.circuit(), cutoff=c,
                     cutoff_type="length", lower=lower, state=p.initial_state)
    _assert_same(o, d)
    if cutoff is None:
        assert abs(d.expectation_value([]) - p.exact_energy) <= 1e-12


@pytest.mark.parametrize("cutoff_type", ["length", "support"])
@pytest.mark.parametrize("n", [5, 40, 100])
def test_random_majorana_circuits(oracle_mod, n, cutoff_type):
    """Weight-2/3/4 generators (odd ones exercise the parity column) over 1..4 words."""
    print(f"case: random majorana n={n} {cutoff_type}", flush=True)
    rng = np.random.default_rng(n)
    gates, params = [], []
    for _ in range(50):
        w = int(rng.choice([2, 3, 4]))
        idx = tuple(sorted(int(i) for i in rng.choice(2 * n, w, replace=False)))
        coeff = {2: 1j, 3: 1j, 4: 1.0}[w]
        gates.append(ExpGate(MajoranaOperator({idx: coeff}, n)))
        params.append(float(rng.uniform(-np.pi, np.pi)))
    circ = Circuit(gates, n, params)
    obs = MajoranaOperator({(0, 1): 0.5j, (2, 3, 4, 5): 0.25, (): 0.5}, n)
    state = list(range(0, n, 2))
    o, d = _run_both(oracle_mod, "majorana", n, obs.terms, circ, cutoff=4, cutoff_type=cutoff_type, lower=1e-5,
                     state=state)
    _assert_same(o, d)


@pytest.mark.parametrize("k", [2, 4])
def test_only_rotate_len_k(oracle_mod, k):
    """Orbital gate (two-pass cos arm, Scan.h:419-433) on the reference's S0_8e8o fixture."""
    print(f"case: only_rotate_len_k={k}", flush=True)
    p = fixtures.load("S0_8e8o_majoranic_c8")
    o, d = _run_both(oracle_mod, "majorana", p.n_modes, p.operator.terms, p.circuit(), cutoff=8,
                     cutoff_type="length", lower=1e-8, state=p.initial_state, only_rot=k)
    _assert_same(o, d)


def test_hubbard_small(oracle_mod):
    print("case: hubbard 3x3 2 steps", flush=True)
    circ, n = workloads.hubbard_lattice(3, 2)
    obs = workloads.number_operator(0, n)
    o, d = _run_both(oracle_mod, "majorana", n, obs.terms, circ, cutoff=8, cutoff_type="length", lower=1e-6,
                     state=circ.initial_state)
    _assert_same(o, d)


def test_lattice_short(oracle_mod):
    """10x10 kicked Ising, 3 steps, theta_h 0.8 (two limbs per plane; bonds straddle qubit 64)."""
    print("case: lattice 10x10 steps=3", flush=True)
    circ = workloads.kicked_ising_lattice(10, 3, theta_h=0.8)
    init = PauliOperator({Pauli("Z", (s,)): 0.01 for s in range(100)}, 100).get_local_operator().terms
    o, d = _run_both(oracle_mod, "pauli", 100, init, circ, cutoff=10, cutoff_type="support", lower=1e-5)
    _assert_same(o, d)


def test_evolved_operator_export(oracle_mod):
    circ = workloads.kicked_ising_chain(12, 3)
    init = workloads.z_observable(12, [6]).get_local_operator().terms
    o, d = _run_both(oracle_mod, "pauli", 12, init, circ, cutoff=6, cutoff_type="support", lower=1e-8)
    eo = o.evolved_operator(1e-12)
    ed = d.evolved_operator([], 1e-12)
    assert eo == ed


@pytest.mark.parametrize("fused", [True, False])
def test_both_pipelines_lattice(oracle_mod, fused):
    """The fused single-sweep pipeline and the per-gate pipeline give the oracle's store; the 840-gate
    circuit crosses several enqueue chunks and store regrowths."""
    print(f"case: lattice 10x10 steps=3 fused={fused}", flush=True)
    circ = workloads.kicked_ising_lattice(10, 3, theta_h=0.4)
    n = 100
    init = PauliOperator({Pauli("Z", (s,)): 0.01 for s in range(n)}, n).get_local_operator().terms
    monos, gcs, pm, _ = expand_monomials(circ.gates, circ.resolved_mapping, n)
    o = oracle_mod.Oracle(init, n, basis="pauli", cutoff=10, cutoff_type="support", lower_atol=1e-5)
    o.propagate(monos, pm, gcs, circ.parameters)
    d = DeviceSimulator(init, 10, [], lower_atol=1e-5, cutoff_type="support", logical_num_modes=n, basis="pauli")
    d.set_pipeline(fused)
    d.propagate(monos, pm, gcs, circ.parameters)
    _assert_same(o, d)


def test_repeated_propagate_calls(oracle_mod):
    """Two propagate calls in a row (the second continues from the first's store and index)."""
    circ = workloads.kicked_ising_chain(16, 2)
    init = workloads.z_observable(16, [8]).get_local_operator().terms
    monos, gcs, pm, _ = expand_monomials(circ.gates, circ.resolved_mapping, 16)
    o = oracle_mod.Oracle(init, 16, basis="pauli", cutoff=6, cutoff_type="support", lower_atol=1e-8)
    d = DeviceSimulator(init, 6, [], lower_atol=1e-8, cutoff_type="support", logical_num_modes=16, basis="pauli")
    for _ in range(2):
        o.propagate(monos, pm, gcs, circ.parameters)
        d.propagate(monos, pm, gcs, circ.parameters)
    _assert_same(o, d)
