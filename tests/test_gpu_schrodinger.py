"""Schrodinger picture (``schrodinger_cutoff``) on the HIP engine vs the oracle.

Contract-immediately propagation is bit-exact with the oracle (rows, coefficients, per-gate counts)
through both pipelines, and its <O> is too: the engine matches the observable's terms against the
rows and sums state . op over those rows in ascending row order, as the reference's dense dot does.
The graph path (build, contract, functional, adjoint gradient) matches the oracle and the reference
fixtures' exact energies and gradients.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import dense
import fixtures
from monoprop_amd import MajoranaPropagator, PauliPropagator, workloads
from monoprop_amd._engine import DeviceSimulator
from monoprop_amd.circuit import Circuit, ExpGate, expand_monomials
from monoprop_amd.majorana import MajoranaOperator
from monoprop_amd.pauli import Pauli, PauliOperator

pytestmark = pytest.mark.gpu


def _make(oracle_mod, basis, n, terms, *, cutoff, cutoff_type, schr, lower=None, upper=None, state=()):
    o = oracle_mod.Oracle(terms, n, basis=basis, cutoff=cutoff, cutoff_type=cutoff_type, lower_atol=lower,
                          upper_atol=upper, initial_state=state, schrodinger_cutoff=schr)
    d = DeviceSimulator(terms, cutoff, list(state), lower_atol=lower, upper_atol=upper, cutoff_type=cutoff_type,
                        logical_num_modes=n, basis=basis, schrodinger_cutoff=schr)
    assert d.schrodinger and o.schrodinger
    return o, d


def _propagate_both(oracle_mod, basis, n, terms, circ, *, only_rot=None, fused=True, **kw):
    monos, gcs, pm, _ = expand_monomials(circ.gates, circ.resolved_mapping, n)
    o, d = _make(oracle_mod, basis, n, terms, **kw)
    if not fused:
        d.set_pipeline(False)
    o.propagate(monos, pm, gcs, circ.parameters, only_rotate_len_k=only_rot)
    d.propagate(monos, pm, gcs, circ.parameters, only_rot)
    return o, d


def _assert_same(o, d):
    wo, co = o.rows()
    wd, cd = d.rows()
    assert wo.shape == wd.shape, (wo.shape, wd.shape)
    assert np.array_equal(wo, wd)
    assert np.array_equal(co, cd), np.max(np.abs(co - cd))
    so, sd = o.gate_stats(), d.gate_stats()
    for a, b in zip(so, sd):
        assert np.array_equal(a, b)
    assert o.expectation() == d.expectation_value([])


@pytest.mark.parametrize("sc", [2, 3])
def test_kicked_ising_chain(oracle_mod, sc):
    """20-qubit chain, 5 steps, cutoff 8, lower_atol 1e-8, observable Z_10, state of sc qubits."""
    print(f"case: schrodinger chain n=20 sc={sc}", flush=True)
    circ = workloads.kicked_ising_chain(20, 5)
    init = workloads.z_observable(20, [10]).get_local_operator().terms
    o, d = _propagate_both(oracle_mod, "pauli", 20, init, circ, cutoff=8, cutoff_type="support", lower=1e-8,
                           schr=2 * sc, state=[3, 10, 11])
    assert o.size() > 5000
    _assert_same(o, d)


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("n", [7, 70, 130])
@pytest.mark.parametrize("lower,upper,cutoff,sc", [(None, None, 3, 2), (1e-3, None, 4, 1), (1e-4, 1e-2, 2, 2)])
def test_random_pauli_circuits(oracle_mod, n, lower, upper, cutoff, sc, fused):
    print(f"case: schrodinger random pauli n={n} lower={lower} upper={upper} cutoff={cutoff} sc={sc} "
          f"fused={fused}", flush=True)
    rng = np.random.default_rng(n * 17 + cutoff)
    gates, params = [], []
    for _ in range(50):
        w = int(rng.integers(1, 4))
        qs = sorted(int(q) for q in rng.choice(n, w, replace=False))
        letters = "".join(str(rng.choice(list("XYZ"))) for _ in qs)
        gates.append(ExpGate(PauliOperator({Pauli(letters, qs): float(rng.choice([1.0, -0.5, 2.0]))}, n)))
        params.append(float(rng.uniform(-np.pi, np.pi)))
    gates.append(ExpGate(PauliOperator({Pauli("ZZ", (0, n - 1)): 1.0}, n)))
    params.append(-np.pi / 4)  # cos 2a = 0: the two-pass arm
    circ = Circuit(gates, n, params)
    obs = PauliOperator({Pauli("Z", (n // 2,)): 1.0, Pauli("XY", (0, n - 1)): -0.3, Pauli("ZZ", (1, 2)): 0.4}, n)
    state = [q for q in range(n) if q % 3 == 0]
    o, d = _propagate_both(oracle_mod, "pauli", n, obs.get_local_operator().terms, circ, cutoff=cutoff,
                           cutoff_type="support", lower=lower, upper=upper, state=state, schr=2 * sc, fused=fused)
    _assert_same(o, d)


@pytest.mark.parametrize("name", fixtures.EXACT_CASES)
@pytest.mark.parametrize("cutoff,lower,sc", [(None, None, None), (4, None, 6), (6, 1e-6, 4)])
def test_majorana_fixtures(oracle_mod, name, cutoff, lower, sc):
    print(f"case: schrodinger fixture {name} cutoff={cutoff} lower={lower} sc={sc}", flush=True)
    p = fixtures.load(name)
    c = cutoff if cutoff is not None else 2 * p.n_modes
    s = sc if sc is not None else 2 * p.n_modes
    o, d = _propagate_both(oracle_mod, "majorana", p.n_modes, p.operator.terms, p.circuit(), cutoff=c,
                           cutoff_type="length", lower=lower, state=p.initial_state, schr=s)
    _assert_same(o, d)
    if cutoff is None:
        assert abs(d.expectation_value([]) - p.exact_energy) <= 1e-12


@pytest.mark.parametrize("cutoff_type", ["length", "support"])
@pytest.mark.parametrize("n", [5, 40, 100])
def test_random_majorana_circuits(oracle_mod, n, cutoff_type):
    print(f"case: schrodinger random majorana n={n} {cutoff_type}", flush=True)
    rng = np.random.default_rng(n + 7)
    gates, params = [], []
    for _ in range(40):
        w = int(rng.choice([2, 3, 4]))
        idx = tuple(sorted(int(i) for i in rng.choice(2 * n, w, replace=False)))
        gates.append(ExpGate(MajoranaOperator({idx: {2: 1j, 3: 1j, 4: 1.0}[w]}, n)))
        params.append(float(rng.uniform(-np.pi, np.pi)))
    circ = Circuit(gates, n, params)
    obs = MajoranaOperator({(0, 1): 0.5j, (2, 3, 4, 5): 0.25, (): 0.5}, n)
    o, d = _propagate_both(oracle_mod, "majorana", n, obs.terms, circ, cutoff=4, cutoff_type=cutoff_type, lower=1e-5,
                           state=list(range(0, n, 2)), schr=4)
    _assert_same(o, d)


@pytest.mark.parametrize("k", [2, 4])
def test_only_rotate_len_k(oracle_mod, k):
    """Orbital gates (two-pass arm) with the reference's Schrodinger setting (cutoff 6, sc 8)."""
    print(f"case: schrodinger only_rotate_len_k={k}", flush=True)
    p = fixtures.load("S0_8e8o_majoranic_c8")
    o, d = _propagate_both(oracle_mod, "majorana", p.n_modes, p.operator.terms, p.circuit(), cutoff=6,
                           cutoff_type="length", lower=1e-8, state=p.initial_state, schr=8, only_rot=k)
    _assert_same(o, d)


def _graph_pair(oracle_mod, basis, n, terms, circ, *, seeded, **kw):
    monos, gcs, pm, gidx = expand_monomials(circ.gates, circ.resolved_mapping, n)
    o, d = _make(oracle_mod, basis, n, terms, **kw)
    o.build_graph(monos, pm, gcs, circ.parameters if seeded else None)
    d.build_graph(monos, pm, gcs, gidx, list(circ.parameters) if seeded else None)
    return o, d


@pytest.mark.parametrize("name", fixtures.EXACT_CASES)
@pytest.mark.parametrize("seeded", [False, True])
def test_graph_fixture_energy_and_gradient(oracle_mod, name, seeded):
    print(f"case: schrodinger graph fixture {name} seeded={seeded}", flush=True)
    p = fixtures.load(name)
    o, d = _graph_pair(oracle_mod, "majorana", p.n_modes, p.operator.terms, p.circuit(), seeded=seeded,
                       cutoff=2 * p.n_modes, cutoff_type="length", state=p.initial_state, schr=2 * p.n_modes)
    wo, co = o.rows()
    wd, cd = d.rows()
    assert np.array_equal(wo, wd)
    assert np.array_equal(co, cd)  # the state at the graph's start, new rows at their scores
    assert np.array_equal(np.asarray(d.contract_partially(p.parameters, False)), o.graph_contract(p.parameters))
    assert abs(d.expectation_value(p.parameters) - p.exact_energy) <= 1e-12
    assert abs(d.expectation_value(p.parameters) - o.graph_ev(p.parameters)) <= 1e-13
    val, grad = d.expectation_value_and_gradient(p.parameters)
    assert abs(val - p.exact_energy) <= 1e-12
    if len(p.exact_gradient):
        assert np.allclose(grad, p.exact_gradient, atol=1e-10)


@pytest.mark.parametrize("name", fixtures.EXACT_CASES)
def test_graph_truncated_gradient(oracle_mod, name):
    """reference test_gradient (schrodinger=True): cutoff 4, schrodinger_cutoff 6, random angles."""
    print(f"case: schrodinger truncated gradient {name}", flush=True)
    p = fixtures.load(name)
    o, d = _graph_pair(oracle_mod, "majorana", p.n_modes, p.operator.terms, p.circuit(), seeded=False, cutoff=4,
                       cutoff_type="length", state=p.initial_state, schr=6)
    x = np.random.default_rng(0).random(len(p.parameters))
    vo, go = o.graph_grad(x)
    vd, gd = d.expectation_value_and_gradient(x)
    assert vo == vd
    assert np.allclose(gd, go, rtol=1e-10, atol=1e-12)
    eps = 1e-6
    fd = np.array([(d.expectation_value(x + eps * e) - d.expectation_value(x - eps * e)) / (2 * eps)
                   for e in np.eye(len(x))])
    assert np.allclose(gd, fd, atol=1e-6)


def test_graph_contract_inplace_then_expectation(oracle_mod):
    print("case: schrodinger contract inplace", flush=True)
    p = fixtures.load("rx_rz_ry_rz_exact")
    o, d = _graph_pair(oracle_mod, "majorana", p.n_modes, p.operator.terms, p.circuit(), seeded=False,
                       cutoff=2 * p.n_modes, cutoff_type="length", state=p.initial_state, schr=2 * p.n_modes)
    state = np.asarray(d.contract_partially(p.parameters, True))
    assert d.graph_layers() == 0
    assert np.array_equal(state, o.graph_contract(p.parameters))
    assert abs(d.expectation_value([]) - p.exact_energy) <= 1e-12


def test_public_api_pauli_schrodinger_cutoff():
    """reference tests/test_pauli.py:47-62 through PauliPropagator: the state's largest weight is sc."""
    print("case: schrodinger public api", flush=True)
    for sc in (3, 4, 5):
        mp = PauliPropagator(PauliOperator({"ZZ": 1.0}, num_qubits=10), initial_state=[], cutoff=4,
                             schrodinger_cutoff=sc)
        assert mp.schrodinger
        op = mp.evolved_operator()
        assert max(len(p.qubits) for p in op.terms) == sc
        assert len(op.terms) == sum(math.comb(10, k) for k in range(sc + 1))


def test_public_api_majorana_matches_heisenberg():
    """A fermionic fixture through MajoranaPropagator: both pictures give the exact energy."""
    print("case: schrodinger majorana public api", flush=True)
    p = fixtures.load("lih_fermionic_spin_exact")
    circ = p.circuit()
    vals = []
    for schr in (None, 2 * p.n_modes):
        mp = MajoranaPropagator(p.operator, p.initial_state, cutoff=2 * p.n_modes, schrodinger_cutoff=schr)
        mp.propagate(circ)
        vals.append(mp.expectation_value())
    assert abs(vals[0] - p.exact_energy) <= 1e-12 and abs(vals[1] - p.exact_energy) <= 1e-12
