"""The thread-parallel restatement (oracle.set_threads) gives the serial oracle's store, coefficients,
per-gate counters and <O> bit for bit, in both pictures and through the graph path."""

from __future__ import annotations

import numpy as np
import pytest

from monoprop_amd import workloads
from monoprop_amd.circuit import expand_monomials


@pytest.mark.parametrize("schr", [None, 2])
def test_threads_match_serial(oracle_mod, schr):
    L, n = 6, 36
    circ = workloads.kicked_ising_lattice(L, 6, theta_h=0.4)
    monos, gcs, pm, _ = expand_monomials(circ.gates, circ.resolved_mapping, n)
    out = []
    for threads in (1, 4):
        oracle_mod.set_threads(threads)
        try:
            o = oracle_mod.Oracle(workloads.mean_z(n).get_local_operator().terms, n, basis="pauli", cutoff=8,
                                  lower_atol=1e-6, schrodinger_cutoff=schr)
            o.propagate(monos, pm, gcs, circ.parameters)
            g = oracle_mod.Oracle(workloads.mean_z(n).get_local_operator().terms, n, basis="pauli", cutoff=6,
                                  lower_atol=1e-6, schrodinger_cutoff=schr)
            g.build_graph(monos[:400], pm[:400], gcs[:400], circ.parameters)
            out.append((o.rows(), o.gate_stats(), o.expectation(), g.rows(), g.graph_ev(circ.parameters)))
        finally:
            oracle_mod.set_threads(1)
    (r1, s1, e1, g1, v1), (r4, s4, e4, g4, v4) = out
    assert r1[0].shape[0] > 20000  # the parallel ranges are used
    assert np.array_equal(r1[0], r4[0]) and np.array_equal(r1[1], r4[1])
    assert all(np.array_equal(a, b) for a, b in zip(s1, s4))
    assert e1 == e4 and v1 == v4
    assert np.array_equal(g1[0], g4[0]) and np.array_equal(g1[1], g4[1])
