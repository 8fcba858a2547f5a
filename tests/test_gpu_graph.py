"""Graph build, replay and adjoint gradient on the HIP engine.

* the store after build_graph and the replayed coefficient vector are bit-exact with the oracle;
* <O> of the replay equals the reference fixtures' exact energies, the gradient their exact gradients;
* the adjoint gradient agrees with the oracle's, with central differences and with the parameter-shift
  rule of the frozen structure (each layer is linear in cos 2a, sin 2a, so
  dE/dtheta = g [E(theta + pi/(4g)) - E(theta - pi/(4g))] when theta drives one layer).
"""

from __future__ import annotations

import numpy as np
import pytest

import fixtures
from monoprop_amd import PauliPropagator, workloads
from monoprop_amd._engine import DeviceSimulator
from monoprop_amd.circuit import expand_monomials

pytestmark = pytest.mark.gpu


def _pair(oracle_mod, basis, n, terms, circ, *, cutoff, cutoff_type, lower=None, upper=None, state=(), seeded=True):
    monos, gcs, pm, gidx = expand_monomials(circ.gates, circ.resolved_mapping, n)
    o = oracle_mod.Oracle(terms, n, basis=basis, cutoff=cutoff, cutoff_type=cutoff_type, lower_atol=lower,
                          upper_atol=upper, initial_state=state)
    o.build_graph(monos, pm, gcs, circ.parameters if seeded else None)
    d = DeviceSimulator(terms, cutoff, list(state), lower_atol=lower, upper_atol=upper, cutoff_type=cutoff_type,
                        logical_num_modes=n, basis=basis)
    d.build_graph(monos, pm, gcs, gidx, list(circ.parameters) if seeded else None)
    return o, d


@pytest.mark.parametrize("name", fixtures.EXACT_CASES)
@pytest.mark.parametrize("seeded", [False, True])
def test_fixture_energy_and_gradient(oracle_mod, name, seeded):
    print(f"case: graph fixture {name} seeded={seeded}", flush=True)
    p = fixtures.load(name)
    o, d = _pair(oracle_mod, "majorana", p.n_modes, p.operator.terms, p.circuit(), cutoff=2 * p.n_modes,
                 cutoff_type="length", state=p.initial_state, seeded=seeded)
    wo, _ = o.rows()
    wd, _ = d.rows()
    assert np.array_equal(wo, wd)
    assert np.array_equal(np.asarray(d.contract_partially(p.parameters, False)), o.graph_contract(p.parameters))
    assert abs(d.expectation_value(p.parameters) - p.exact_energy) <= 1e-12
    val, grad = d.expectation_value_and_gradient(p.parameters)
    assert abs(val - p.exact_energy) <= 1e-12
    if len(p.exact_gradient):
        assert np.allclose(grad, p.exact_gradient, atol=1e-10)


@pytest.mark.parametrize("name", fixtures.EXACT_CASES)
def test_truncated_gradient(oracle_mod, name):
    """cutoff 4 at random angles: gradient vs oracle, central differences and parameter shift."""
    print(f"case: truncated gradient {name}", flush=True)
    p = fixtures.load(name)
    o, d = _pair(oracle_mod, "majorana", p.n_modes, p.operator.terms, p.circuit(), cutoff=4, cutoff_type="length",
                 lower=1e-6, state=p.initial_state)
    x = np.random.default_rng(0).random(len(p.parameters))
    vo, go = o.graph_grad(x)
    vd, gd = d.expectation_value_and_gradient(x)
    assert vo == vd
    assert np.allclose(gd, go, rtol=1e-10, atol=1e-12)
    eps = 1e-6
    fd = np.array([(d.expectation_value(x + eps * e) - d.expectation_value(x - eps * e)) / (2 * eps)
                   for e in np.eye(len(x))])
    assert np.allclose(gd, fd, atol=1e-6)


def test_ring_brickwork_small(oracle_mod):
    """64-qubit brickwork ring, 2 layers, cutoff 8, lower_atol 1e-6, 384 independent angles."""
    print("case: ring brickwork 2 layers", flush=True)
    circ = workloads.brickwork_ring(64, 2, seed=0)
    terms = workloads.ring_zz_mean(64).get_local_operator().terms
    o, d = _pair(oracle_mod, "pauli", 64, terms, circ, cutoff=8, cutoff_type="support", lower=1e-6)
    wo, _ = o.rows()
    wd, _ = d.rows()
    assert np.array_equal(wo, wd)
    theta = np.asarray(circ.parameters)
    assert np.array_equal(np.asarray(d.contract_partially(theta, False)), o.graph_contract(theta))
    vo, go = o.graph_grad(theta)
    vd, gd = d.expectation_value_and_gradient(theta)
    assert vo == vd
    assert np.allclose(gd, go, rtol=1e-9, atol=1e-13)
    # parameter shift on 16 sampled angles (every gate has g = 1 and its own angle)
    idx = np.random.default_rng(1).choice(len(theta), 16, replace=False)
    for i in idx:
        e = np.zeros_like(theta)
        e[i] = np.pi / 4
        ps = d.expectation_value(theta + e) - d.expectation_value(theta - e)
        assert abs(ps - gd[i]) <= 1e-11


def test_contract_inplace_then_expectation(oracle_mod):
    circ = workloads.kicked_ising_chain(12, 2)
    prop = PauliPropagator(workloads.z_observable(12, [6]), [], cutoff=6, lower_atol=1e-8)
    prop.build_graph(circ)
    ev_graph = prop.expectation_value(circ.parameters)
    prop.contract_partially(circ.parameters, inplace=True)
    assert prop.graph_layers == 0
    assert abs(prop.expectation_value() - ev_graph) <= 1e-14


def test_graph_matches_propagate(oracle_mod):
    """Seeded build then replay reproduces the contract-immediately <O> (reference
    test_coeff_trunc_build_graph_and_inplace_equiv) and the same term count."""
    circ = workloads.kicked_ising_chain(20, 3)
    a = PauliPropagator(workloads.z_observable(20, [10]), [], cutoff=8, lower_atol=1e-8)
    a.propagate(circ)
    b = PauliPropagator(workloads.z_observable(20, [10]), [], cutoff=8, lower_atol=1e-8)
    b.build_graph(circ)
    assert a.size() == b.size()
    assert abs(a.expectation_value() - b.expectation_value(circ.parameters)) <= 1e-12


def test_torch_autograd_matches_adjoint():
    """theta.grad through monoprop_amd.autograd equals the engine's adjoint gradient (times the
    incoming gradient), and the forward value equals the replayed <O>."""
    import torch

    from monoprop_amd.autograd import expectation

    circ = workloads.brickwork_ring(16, 2, seed=3)
    prop = PauliPropagator(workloads.ring_zz_mean(16), [], cutoff=6, lower_atol=1e-6)
    prop.build_graph(circ)
    theta0 = np.asarray(circ.parameters)
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    ev = expectation(prop, theta)
    (3.0 * ev).backward()
    v, g = prop.expectation_value_and_gradient(theta0)
    assert abs(ev.item() - v) <= 1e-14
    assert np.allclose(theta.grad.numpy(), 3.0 * np.asarray(g), rtol=0, atol=1e-14)
    with pytest.raises(ValueError):
        expectation(prop, theta[:-1])


@pytest.mark.parametrize("schr", [None, 6])
def test_checkpointed_gradient_independent_of_segments(oracle_mod, schr):
    """The adjoint replays segments again from coefficient-vector checkpoints: the value and every
    gradient entry are the same bits for one segment (a budget that holds every snapshot) and for
    the most segments (a zero budget), agree with the oracle's adjoint, and a gradient that follows a
    kept checkpointed forward at the same parameters gives those bits too."""
    print(f"case: checkpointed gradient schr={schr}", flush=True)
    p = fixtures.load("lih_fermionic_spin_exact")  # 72 graph layers
    monos, gcs, pm, gidx = expand_monomials(p.circuit().gates, p.circuit().resolved_mapping, p.n_modes)
    d = DeviceSimulator(p.operator.terms, 6, list(p.initial_state), lower_atol=1e-6, cutoff_type="length",
                        logical_num_modes=p.n_modes, basis="majorana", schrodinger_cutoff=schr)
    d.build_graph(monos, pm, gcs, gidx, None)
    x = list(np.random.default_rng(7).random(len(p.parameters)))
    d.set_gradient_checkpointing(False, 1 << 40)
    v1, g1 = d.expectation_value_and_gradient(x)
    assert d.gradient_segments()["segments"] == 1
    d.set_gradient_checkpointing(False, 0)
    v2, g2 = d.expectation_value_and_gradient(x)
    segs = d.gradient_segments()
    assert segs["segments"] > 16 and segs["checkpoint_bytes"] > 0
    assert v1 == v2 and np.array_equal(np.asarray(g1), np.asarray(g2))
    assert v1 == d.expectation_value(x)
    d.set_gradient_checkpointing(True, 0)
    assert d.expectation_value(x) == v1
    v3, g3 = d.expectation_value_and_gradient(x)
    assert v3 == v1 and np.array_equal(np.asarray(g3), np.asarray(g1))
    y = list(np.asarray(x) + 0.25)
    assert d.expectation_value(y) != v1  # a kept forward at other parameters is not reused for x
    v4, g4 = d.expectation_value_and_gradient(x)
    assert v4 == v1 and np.array_equal(np.asarray(g4), np.asarray(g1))
    o = oracle_mod.Oracle(p.operator.terms, p.n_modes, basis="majorana", cutoff=6, cutoff_type="length",
                          lower_atol=1e-6, initial_state=p.initial_state, schrodinger_cutoff=schr)
    o.build_graph(monos, pm, gcs, None)
    vo, go = o.graph_grad(np.asarray(x))
    assert vo == v1
    assert np.allclose(g1, go, rtol=1e-10, atol=1e-12)


def _build(basis, n, terms, circ, pipeline_fused, *, cutoff, cutoff_type, lower, state=(), schr=None, seeded=True):
    monos, gcs, pm, gidx = expand_monomials(circ.gates, circ.resolved_mapping, n)
    d = DeviceSimulator(terms, cutoff, list(state), lower_atol=lower, cutoff_type=cutoff_type, logical_num_modes=n,
                        basis=basis, schrodinger_cutoff=schr)
    d.set_pipeline(pipeline_fused)
    d.build_graph(monos, pm, gcs, gidx, list(circ.parameters) if seeded else None)
    return d


@pytest.mark.parametrize("case", ["ring64", "lattice5", "hubbard3", "ring8_schr", "ring64_structural"])
def test_graph_build_pipelines_agree(oracle_mod, case):
    """A graph built on the fused kernels (the default without an orbital cap) and one built by the
    per-gate pipeline hold the same store, the same layer records (rotation and recomputed cos-only
    counts), replay to the same coefficient vector bit for bit, and give the oracle's replayed <O>; their
    gradients agree to rounding (the records' order inside a layer differs)."""
    print(f"case: graph build pipelines {case}", flush=True)
    schr, seeded = None, True
    if case == "ring64":
        circ, n, basis, ct, cut, low = workloads.brickwork_ring(64, 3, seed=5), 64, "pauli", "support", 8, 1e-6
        terms, state = workloads.ring_zz_mean(64).get_local_operator().terms, []
    elif case == "ring64_structural":
        circ, n, basis, ct, cut, low = workloads.brickwork_ring(64, 1, seed=2), 64, "pauli", "support", 4, None
        terms, state, seeded = workloads.ring_zz_mean(64).get_local_operator().terms, [], False
    elif case == "lattice5":
        circ, n, basis, ct, cut, low = workloads.kicked_ising_lattice(5, 3, theta_h=0.4), 25, "pauli", "support", 8, 1e-5
        terms, state = workloads.mean_z(25).get_local_operator().terms, []
    elif case == "hubbard3":
        circ, n = workloads.hubbard_lattice(3, 2)
        basis, ct, cut, low = "majorana", "length", 6, 1e-6
        terms, state = workloads.number_operator(0, n).terms, list(circ.initial_state)
    else:
        circ, n, basis, ct, cut, low = workloads.brickwork_ring(8, 2, seed=1), 8, "pauli", "support", 4, 1e-8
        terms, state, schr = workloads.ring_zz_mean(8).get_local_operator().terms, [0, 5], 4
    a = _build(basis, n, terms, circ, True, cutoff=cut, cutoff_type=ct, lower=low, state=state, schr=schr,
               seeded=seeded)
    b = _build(basis, n, terms, circ, False, cutoff=cut, cutoff_type=ct, lower=low, state=state, schr=schr,
               seeded=seeded)
    wa, ca = a.rows()
    wb, cb = b.rows()
    assert np.array_equal(wa, wb) and np.array_equal(ca, cb)
    assert a.graph_size() == b.graph_size() and a.graph_layers() == b.graph_layers()
    theta = np.asarray(circ.parameters, dtype=np.float64) * 0.9 + 0.05
    assert np.array_equal(np.asarray(a.contract_partially(theta, False)), np.asarray(b.contract_partially(theta, False)))
    va, ga = a.expectation_value_and_gradient(theta)
    vb, gb = b.expectation_value_and_gradient(theta)
    assert va == vb
    assert np.allclose(ga, gb, rtol=1e-10, atol=1e-14)
    if schr is None:
        monos, gcs, pm, _ = expand_monomials(circ.gates, circ.resolved_mapping, n)
        o = oracle_mod.Oracle(terms, n, basis=basis, cutoff=cut, cutoff_type=ct, lower_atol=low, initial_state=state)
        o.build_graph(monos, pm, gcs, circ.parameters if seeded else None)
        vo, go = o.graph_grad(theta)
        assert vo == va
        assert np.allclose(ga, go, rtol=1e-9, atol=1e-13)
