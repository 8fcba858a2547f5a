"""Tight term stores on the HIP engine: a lone propagation whose store may not grow to a gate's 2N
worst case (a store limit below twice the live terms) is bounded gate by gate by the gate's own
anticommuting terms, counted on the device, and still hashes to the oracle's golden digests with <O>
to the bit; the store never holds more rows than the limit; a graph built on a tight store replays like
one built without a limit; a limit the propagation cannot fit in raises the engine's RuntimeError."""

from __future__ import annotations

import importlib.util
import json
from pathlib import Path

import numpy as np
import pytest

from monoprop_amd._engine import DeviceSimulator

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
_spec = importlib.util.spec_from_file_location("make_golden", GOLDEN / "make_golden.py")
make_golden = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(make_golden)
VECTORS = json.loads((GOLDEN / "oracle_vectors.json").read_text())


def _sim(case):
    return DeviceSimulator(case["obs"], case["cutoff"], case["state"], lower_atol=case["lower"], upper_atol=case["upper"],
                           cutoff_type=case["cutoff_type"], logical_num_modes=case["n"], basis=case["basis"],
                           schrodinger_cutoff=case["schr"])


def _capacity_rows(d):
    """Rows of store capacity: the live terms plus the slack rows (d_terms_slack_bytes counts the
    capacity's unused rows; the arrays themselves may sit in larger cached device blocks)."""
    words = d.rows()[0].shape[1]
    return d.size() + d.operator_memory_breakdown()["d_terms_slack_bytes"] // (8 * words)


@pytest.mark.parametrize("name,factor", [("hubbard6_3steps", 1.4), ("hubbard6_2steps", 1.4), ("lattice10_full_h04", 1.25)])
def test_tight_store_matches_golden(name, factor):
    print(f"case: tight store {name} x{factor}", flush=True)
    case = make_golden.cases()[name]
    want = VECTORS[name]
    circ = make_golden.build_circuit(case["circuit"])
    monos, gcs, pm, _ = make_golden.layers(circ)
    d = _sim(case)
    limit = int(want["terms"] * factor)  # capacities are multiples of 16384 rows: limits well above that
    d.set_store_limit(limit)
    d.propagate(monos, pm, gcs, make_golden.parameters(case, circ))
    w, c = d.rows()
    assert d.size() == want["terms"]
    assert make_golden.digest(np.ascontiguousarray(w, dtype=np.uint64)) == want["rows_sha256"]
    assert make_golden.digest(np.ascontiguousarray(c, dtype=np.float64)) == want["coeffs_sha256"]
    assert d.expectation_value([]).hex() == want["expectation_hex"]
    assert _capacity_rows(d) <= limit + 16384


def test_tight_store_limit_too_small_raises():
    print("case: tight store, limit below the final size", flush=True)
    case = make_golden.cases()["hubbard6_3steps"]
    circ = make_golden.build_circuit(case["circuit"])
    monos, gcs, pm, _ = make_golden.layers(circ)
    d = _sim(case)
    d.set_store_limit(VECTORS["hubbard6_3steps"]["terms"] // 2)
    with pytest.raises(RuntimeError, match="term store limit"):
        d.propagate(monos, pm, gcs, make_golden.parameters(case, circ))


def test_tight_store_graph_build():
    """A graph build whose store is tight keeps the same layers and replays to the same vector."""
    print("case: tight store graph build, hubbard 6x6 2 steps", flush=True)
    case = make_golden.cases()["hubbard6_2steps"]
    circ = make_golden.build_circuit(case["circuit"])
    monos, gcs, pm, gidx = make_golden.layers(circ)
    params = make_golden.parameters(case, circ)
    ref, d = _sim(case), _sim(case)
    ref.build_graph(monos, pm, gcs, gidx, params)
    d.set_store_limit(int(ref.size() * 1.4))
    d.build_graph(monos, pm, gcs, gidx, params)
    assert d.size() == ref.size() and d.graph_size() == ref.graph_size()
    assert np.array_equal(d.rows()[0], ref.rows()[0])
    theta = np.asarray(params) * 0.7
    assert np.array_equal(np.asarray(d.contract_partially(theta, False)), np.asarray(ref.contract_partially(theta, False)))
