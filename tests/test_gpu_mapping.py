"""Re-wiring a stored graph's parameter mapping (reference MonomialPropagator.inl:809-852,
monomial_propagator.py:314-335), checked by what the re-wiring means: each layer rotates by
theta[new index] * g, so the re-wired graph at y equals the original at the x that gives every layer
the same angle, to the bit; its gradient is the original's summed over the tied parameters.  A
functional created before the re-wiring keeps the mapping it was built with."""

from __future__ import annotations

import numpy as np
import pytest

from monoprop_amd import MajoranaPropagator, PauliPropagator, workloads

pytestmark = pytest.mark.gpu


def test_per_layer_mapping_ties_every_layer():
    print("case: mapping per layer", flush=True)
    circ = workloads.brickwork_ring(8, 2, seed=3)
    obs = workloads.ring_zz_mean(8)
    mp = PauliPropagator(obs, [], cutoff=6, lower_atol=1e-8)
    mp.build_graph(circ)
    n_layers = mp.graph_layers
    assert n_layers == len(circ.gates) == mp.n_parameters
    x = np.asarray(circ.parameters)
    old_fn = mp.expval_functional()
    e0, g0 = mp.expval_and_grad(x)
    # every layer on parameter 0 or 1 by parity of its optimizer position
    mapping = [i % 2 for i in range(n_layers)]
    mp.parameter_mapping = mapping
    assert mp.parameter_mapping == mapping and mp.n_parameters == 2
    y = np.array([0.37, -1.1])
    xt = np.array([y[m] for m in mapping])
    e1, g1 = mp.expval_and_grad(y)
    assert e1 == mp_expval_at(obs, circ, xt)
    gt = grad_at(obs, circ, xt)
    tied = [gt[np.array(mapping) == k].sum() for k in (0, 1)]
    assert np.allclose(g1, tied, rtol=1e-10, atol=1e-13)
    # the functional built before keeps the original mapping
    assert old_fn(x) == e0
    assert mp.expval_functional()(y) == e1


def mp_expval_at(obs, circ, x):
    mp = PauliPropagator(obs, [], cutoff=6, lower_atol=1e-8)
    mp.build_graph(circ)
    return mp.expval(x)


def grad_at(obs, circ, x):
    mp = PauliPropagator(obs, [], cutoff=6, lower_atol=1e-8)
    mp.build_graph(circ)
    return np.asarray(mp.expval_and_grad(x)[1])


def test_per_gate_mapping_keeps_a_gates_layers_together():
    """Hubbard hopping gates have two monomials (two layers) each: a per-gate mapping moves both."""
    print("case: mapping per gate", flush=True)
    circ, n_modes = workloads.hubbard_lattice(2, 1)
    obs = workloads.number_operator(0, n_modes)
    mp = MajoranaPropagator(obs, circ.initial_state, cutoff=6, lower_atol=1e-10)
    mp.build_graph(circ)
    n_gates = len(circ.gates)
    assert mp.n_gates == n_gates < mp.graph_layers
    gate_map = [g % 3 for g in range(n_gates)]
    mp.parameter_mapping = gate_map
    assert mp.n_parameters == 3
    y = np.array([0.21, -0.4, 0.9])
    ref = MajoranaPropagator(obs, circ.initial_state, cutoff=6, lower_atol=1e-10)
    ref.build_graph(circ)
    x = np.array([y[gate_map[g]] for g in range(n_gates)])  # gate g is driven by its own parameter g
    assert mp.expval(y) == ref.expval(x)
    with pytest.raises((ValueError, RuntimeError)):
        mp.parameter_mapping = [0, 1]  # neither per layer nor per gate


def test_deepcopy_is_independent(oracle_mod):
    """copy.deepcopy gives an independent engine (the reference's __deepcopy__): same store, graph and
    answers, and work on one leaves the other untouched."""
    import copy

    from monoprop_amd.circuit import expand_monomials

    print("case: deepcopy", flush=True)
    n = 20
    circ = workloads.kicked_ising_chain(n, 3)
    obs = workloads.z_observable(n, [10])
    a = PauliPropagator(obs, [], cutoff=8, lower_atol=1e-8)
    a.propagate(circ)
    b = copy.deepcopy(a)
    wa, ca = a._simulator.rows()
    wb, cb = b._simulator.rows()
    assert np.array_equal(wa, wb) and np.array_equal(ca, cb) and a.expval() == b.expval()
    b.propagate(circ)  # three more steps on the copy only (Heisenberg: prepended to the circuit)
    assert a.size() == len(ca) and np.array_equal(a._simulator.rows()[1], ca)
    monos, gcs, pm, _ = expand_monomials(circ.gates, circ.resolved_mapping, n)
    o = oracle_mod.Oracle(obs.get_local_operator().terms, n, basis="pauli", cutoff=8, lower_atol=1e-8)
    o.propagate(monos, pm, gcs, circ.parameters)
    o.propagate(monos, pm, gcs, circ.parameters)
    wo, co = o.rows()
    wb, cb = b._simulator.rows()
    assert np.array_equal(wo, wb) and np.array_equal(co, cb)
    # a stored graph travels with the copy
    g = PauliPropagator(workloads.ring_zz_mean(8), [], cutoff=6, lower_atol=1e-8)
    g.build_graph(workloads.brickwork_ring(8, 2, seed=3))
    x = np.asarray(workloads.brickwork_ring(8, 2, seed=3).parameters)
    h = copy.deepcopy(g)
    assert h.graph_layers == g.graph_layers and h.expval_and_grad(x)[0] == g.expval_and_grad(x)[0]
    assert np.array_equal(h.expval_and_grad(x)[1], g.expval_and_grad(x)[1])
    del g
    assert h.expval(x) == h.expval(x)
