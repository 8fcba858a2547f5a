"""Seeded random circuits, HIP engine against the oracle on every path at once: the fused propagation
(store, coefficients, per-gate counters and <O> to the bit), the graph build on the fused kernels
(store and layers, the replayed coefficient vector to the bit, <O> to the bit, the adjoint gradient to
rounding) and, for Schrodinger cases, the evolved state.  Each seed draws the basis, width, generator
weights and coefficients, angles, truncation (lower / upper atol, cutoff) and initial state.  The
Heisenberg circuits are also propagated as batches of several parameter vectors, member by member
against the oracle."""

from __future__ import annotations

import numpy as np
import pytest

from monoprop_amd._engine import DeviceSimulator
from monoprop_amd.circuit import Circuit, ExpGate, expand_monomials
from monoprop_amd.majorana import MajoranaOperator
from monoprop_amd.pauli import Pauli, PauliOperator

pytestmark = pytest.mark.gpu


def _case(seed):
    rng = np.random.default_rng(1000 + seed)
    pauli = bool(seed % 2 == 0)
    n = int(rng.choice([9, 70, 130] if pauli else [6, 40, 100]))
    gates, params = [], []
    for _ in range(int(rng.integers(30, 150))):
        if pauli:
            w = int(rng.integers(1, 4))
            qs = sorted(int(q) for q in rng.choice(n, w, replace=False))
            letters = "".join(str(rng.choice(list("XYZ"))) for _ in qs)
            gates.append(ExpGate(PauliOperator({Pauli(letters, qs): float(rng.choice([1.0, -0.5, 1.5]))}, n)))
        else:
            w = int(rng.choice([2, 3, 4]))
            idx = tuple(sorted(int(i) for i in rng.choice(2 * n, w, replace=False)))
            gates.append(ExpGate(MajoranaOperator({idx: {2: 1j, 3: 1j, 4: 1.0}[w] * float(rng.choice([1.0, -0.7]))}, n)))
        params.append(float(rng.uniform(-np.pi, np.pi)))
    circ = Circuit(gates, n, params)
    if pauli:
        a, b = (int(q) for q in rng.choice(n, 2, replace=False))
        obs = PauliOperator({Pauli("Z", (a,)): 1.0, Pauli("XY", tuple(sorted((a, b)))): -0.4}, n).get_local_operator().terms
        cutoff_type, cutoff = "support", int(rng.integers(2, 5))
    else:
        obs = MajoranaOperator({(0, 1): 0.5j, (2, 3, 4, 5): 0.25, (): 0.5}, n).terms
        cutoff_type, cutoff = str(rng.choice(["length", "support"])), int(rng.integers(3, 7))
    lower = [None, 1e-6, 1e-3][int(rng.integers(0, 3))]
    upper = [None, 1e-2, 0.3][int(rng.integers(0, 3))]
    if lower is not None and upper is not None and upper < lower:
        upper = None
    state = sorted(int(q) for q in rng.choice(n, int(rng.integers(0, n // 2 + 1)), replace=False))
    schr = 4 if seed % 5 == 4 else None
    return dict(basis="pauli" if pauli else "majorana", n=n, circ=circ, obs=obs, cutoff=cutoff,
                cutoff_type=cutoff_type, lower=lower, upper=upper, state=state, schr=schr)


def _pair(oracle_mod, c):
    o = oracle_mod.Oracle(c["obs"], c["n"], basis=c["basis"], cutoff=c["cutoff"], cutoff_type=c["cutoff_type"],
                          lower_atol=c["lower"], upper_atol=c["upper"], initial_state=c["state"],
                          schrodinger_cutoff=c["schr"])
    d = DeviceSimulator(c["obs"], c["cutoff"], list(c["state"]), lower_atol=c["lower"], upper_atol=c["upper"],
                        cutoff_type=c["cutoff_type"], logical_num_modes=c["n"], basis=c["basis"],
                        schrodinger_cutoff=c["schr"])
    return o, d


@pytest.mark.parametrize("seed", range(40))
def test_random_circuit_all_paths(oracle_mod, seed):
    c = _case(seed)
    print(f"case: random seed={seed} {c['basis']} n={c['n']} gates={len(c['circ'].gates)} cutoff={c['cutoff']} "
          f"{c['cutoff_type']} lower={c['lower']} upper={c['upper']} schr={c['schr']}", flush=True)
    circ, n = c["circ"], c["n"]
    monos, gcs, pm, gidx = expand_monomials(circ.gates, circ.resolved_mapping, n)
    theta = list(circ.parameters)
    # propagate
    o, d = _pair(oracle_mod, c)
    o.propagate(monos, pm, gcs, theta)
    d.propagate(monos, pm, gcs, theta)
    wo, co = o.rows()
    wd, cd = d.rows()
    assert np.array_equal(wo, wd) and np.array_equal(co, cd)
    for x, y in zip(o.gate_stats(), d.gate_stats()):
        assert np.array_equal(x, y)
    assert o.expectation() == d.expectation_value([])
    # graph build at the same angles, then replay and adjoint at other angles
    o, d = _pair(oracle_mod, c)
    o.build_graph(monos, pm, gcs, theta)
    d.build_graph(monos, pm, gcs, gidx, theta)
    assert np.array_equal(o.rows()[0], d.rows()[0])
    assert d.graph_layers() == o.graph_layers()
    x = np.asarray(theta) * 0.8 + 0.1
    if c["schr"] is None:
        assert np.array_equal(o.graph_contract(x), np.asarray(d.contract_partially(x, False)))
    vo, go = o.graph_grad(x)
    vd, gd = d.expectation_value_and_gradient(x)
    assert vo == vd
    assert np.allclose(gd, go, rtol=1e-9, atol=1e-13)


@pytest.mark.parametrize("seed", [s for s in range(40) if s % 5 != 4][:30])
def test_random_circuit_batched(oracle_mod, seed):
    """The same circuits propagated as one batch of 2-5 members, each at its own parameter vector: every
    member's rows, coefficients and <O> equal its own oracle propagation to the bit."""
    from monoprop_amd._engine import BatchSimulator

    c = _case(seed)
    rng = np.random.default_rng(7000 + seed)
    circ, n = c["circ"], c["n"]
    k = int(rng.integers(2, 6))
    print(f"case: random batched seed={seed} {c['basis']} n={n} members={k}", flush=True)
    monos, gcs, pm, _ = expand_monomials(circ.gates, circ.resolved_mapping, n)
    params = np.asarray([np.asarray(circ.parameters) * float(rng.uniform(0.5, 1.5)) + float(rng.uniform(-0.3, 0.3))
                         for _ in range(k)], dtype=np.float64)
    d = BatchSimulator([c["obs"]] * k, c["cutoff"], list(c["state"]), lower_atol=c["lower"], upper_atol=c["upper"],
                       cutoff_type=c["cutoff_type"], logical_num_modes=n, basis=c["basis"])
    d.propagate(monos, pm, gcs, params)
    w, co, member = d.rows()
    evs = d.expectation_values()
    for q in range(k):
        o = oracle_mod.Oracle(c["obs"], n, basis=c["basis"], cutoff=c["cutoff"], cutoff_type=c["cutoff_type"],
                              lower_atol=c["lower"], upper_atol=c["upper"], initial_state=c["state"])
        o.propagate(monos, pm, gcs, list(params[q]))
        wo, cq = o.rows()
        sel = member == q
        assert np.array_equal(wo, w[sel]) and np.array_equal(cq, co[sel])
        assert o.expectation() == evs[q]
