"""Oracle vs the reference's exact-energy fixtures (reference tests/test_infinite_cutoff.py:42-70).

At infinite cutoff the propagated expectation must equal the fixture's exact energy to 1e-12, through
both the contract-immediately path and the graph replay; the graph gradient must match the exact
gradient the fixture stores and central finite differences.
"""

from __future__ import annotations

import numpy as np
import pytest

import fixtures
from monoprop_amd.circuit import expand_monomials


def _oracle(oracle_mod, p, **kw):
    kw.setdefault("cutoff", 2 * p.n_modes)
    return oracle_mod.Oracle(p.operator.terms, p.n_modes, basis="majorana", cutoff_type="length",
                             initial_state=p.initial_state, **kw)


def _layers(p):
    c = p.circuit()
    return expand_monomials(c.gates, c.resolved_mapping, p.n_modes)


@pytest.mark.parametrize("name", fixtures.EXACT_CASES)
def test_infinite_cutoff_propagate(oracle_mod, name):
    p = fixtures.load(name)
    monos, gcs, pm, _ = _layers(p)
    o = _oracle(oracle_mod, p)
    o.propagate(monos, pm, gcs, p.parameters)
    assert abs(o.expectation() - p.exact_energy) <= 1e-12


@pytest.mark.parametrize("name", fixtures.EXACT_CASES)
def test_infinite_cutoff_immediate_with_atol(oracle_mod, name):
    """reference test_immediate_contraction: lower_atol=1e-15 still reproduces the exact energy."""
    p = fixtures.load(name)
    monos, gcs, pm, _ = _layers(p)
    o = _oracle(oracle_mod, p, lower_atol=1e-15)
    o.propagate(monos, pm, gcs, p.parameters)
    # the reference asserts np.isclose(atol=1e-12), i.e. 1e-12 + 1e-5 relative; the 1e-15 emission gate
    # drops ~1e-12 of weight on LiH, so the absolute bound here is 1e-11
    assert abs(o.expectation() - p.exact_energy) <= 1e-11


@pytest.mark.parametrize("name", fixtures.EXACT_CASES)
@pytest.mark.parametrize("seeded", [False, True])
def test_infinite_cutoff_graph(oracle_mod, name, seeded):
    p = fixtures.load(name)
    monos, gcs, pm, _ = _layers(p)
    o = _oracle(oracle_mod, p)
    o.build_graph(monos, pm, gcs, p.parameters if seeded else None)
    assert abs(o.graph_ev(p.parameters) - p.exact_energy) <= 1e-12
    val, grad = o.graph_grad(p.parameters)
    assert abs(val - p.exact_energy) <= 1e-12
    if len(p.exact_gradient):
        assert np.allclose(grad, p.exact_gradient, atol=1e-10)


@pytest.mark.parametrize("name", fixtures.EXACT_CASES)
def test_truncated_gradient_vs_fd(oracle_mod, name):
    """reference test_gradient: cutoff 4, random angles, gradient vs central differences (1e-6)."""
    p = fixtures.load(name)
    monos, gcs, pm, _ = _layers(p)
    o = _oracle(oracle_mod, p, cutoff=4)
    o.build_graph(monos, pm, gcs, None)
    rng = np.random.default_rng(0)
    x = rng.random(len(p.parameters))
    val, grad = o.graph_grad(x)
    assert abs(val - o.graph_ev(x)) <= 1e-13
    eps = 1e-6
    fd = np.zeros_like(x)
    for i in range(len(x)):
        xp, xm = x.copy(), x.copy()
        xp[i] += eps
        xm[i] -= eps
        fd[i] = (o.graph_ev(xp) - o.graph_ev(xm)) / (2 * eps)
    assert np.allclose(grad, fd, atol=1e-6)


@pytest.mark.parametrize("name", fixtures.EXACT_CASES)
@pytest.mark.parametrize("upper", [1.0, 1e-2, 1e-4])
@pytest.mark.parametrize("lower", [1e-5, 1e-6])
@pytest.mark.parametrize("cutoff", [4, 6])
def test_seeded_graph_matches_propagate(oracle_mod, name, upper, lower, cutoff):
    """reference test_coeff_trunc_build_graph_and_inplace_equiv: same size, <O> within 1e-12."""
    p = fixtures.load(name)
    monos, gcs, pm, _ = _layers(p)
    a = _oracle(oracle_mod, p, cutoff=cutoff, lower_atol=lower, upper_atol=upper)
    a.propagate(monos, pm, gcs, p.parameters)
    b = _oracle(oracle_mod, p, cutoff=cutoff, lower_atol=lower, upper_atol=upper)
    b.build_graph(monos, pm, gcs, p.parameters)
    assert a.size() == b.size()
    assert abs(a.expectation() - b.graph_ev(p.parameters)) <= 1e-12


def test_ring_gradient_parameter_shift(oracle_mod):
    """64-qubit brickwork ring, 2 layers: the oracle adjoint equals the frozen-structure parameter shift
    dE/dtheta = E(theta + pi/4) - E(theta - pi/4) (every gate has g = 1 and its own angle)."""
    from monoprop_amd import workloads

    circ = workloads.brickwork_ring(64, 2, seed=0)
    monos, gcs, pm, _ = expand_monomials(circ.gates, circ.resolved_mapping, 64)
    o = oracle_mod.Oracle(workloads.ring_zz_mean(64).get_local_operator().terms, 64, basis="pauli", cutoff=8,
                          lower_atol=1e-6)
    o.build_graph(monos, pm, gcs, circ.parameters)
    theta = np.asarray(circ.parameters)
    val, grad = o.graph_grad(theta)
    assert abs(val - o.graph_ev(theta)) <= 1e-14
    for i in np.random.default_rng(1).choice(len(theta), 12, replace=False):
        e = np.zeros_like(theta)
        e[i] = np.pi / 4
        assert abs((o.graph_ev(theta + e) - o.graph_ev(theta - e)) - grad[i]) <= 1e-12
