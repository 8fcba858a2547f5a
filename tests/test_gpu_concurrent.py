"""Propagators driven from several host threads at once (monoprop_amd.concurrent): every engine has
its own stream, and each one's store, FP64 coefficients, per-gate counters and <O> are those of a
lone run, to the bit (the committed golden digests of the 10x10 kicked-Ising lattice at full depth,
20 Trotter steps, theta_h 0.1 / 0.4 / 0.8, and of the 64-qubit ring over 12 layers)."""

from __future__ import annotations

import importlib.util
import json
from pathlib import Path

import numpy as np
import pytest

from monoprop_amd._engine import DeviceSimulator
from monoprop_amd.concurrent import run_concurrently

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
_spec = importlib.util.spec_from_file_location("make_golden", GOLDEN / "make_golden.py")
make_golden = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(make_golden)
VECTORS = json.loads((GOLDEN / "oracle_vectors.json").read_text())
NAMES = ["lattice10_full_h01", "lattice10_full_h04", "lattice10_full_h08", "ring64_12layers"]


def _run(name):
    case = make_golden.cases()[name]
    circ = make_golden.build_circuit(case["circuit"])
    monos, gcs, pm, _ = make_golden.layers(circ)
    d = DeviceSimulator(case["obs"], case["cutoff"], case["state"], lower_atol=case["lower"],
                        upper_atol=case["upper"], cutoff_type=case["cutoff_type"], logical_num_modes=case["n"],
                        basis=case["basis"], schrodinger_cutoff=case["schr"])
    d.propagate(monos, pm, gcs, make_golden.parameters(case, circ))
    w, c = d.rows()
    n, a, e, i = d.gate_stats()
    return (d.size(), make_golden.digest(np.ascontiguousarray(w, dtype=np.uint64)),
            make_golden.digest(np.ascontiguousarray(c, dtype=np.float64)), make_golden.stats_digest(n, a, e, i),
            d.expectation_value([]).hex())


@pytest.mark.parametrize("workers", [None, 2])
def test_concurrent_propagations_match_golden(workers):
    print(f"case: concurrent propagations, workers={workers}", flush=True)
    got = run_concurrently([lambda name=name: _run(name) for name in NAMES], max_workers=workers)
    for name, (terms, rows, coeffs, stats, ev) in zip(NAMES, got):
        want = VECTORS[name]
        assert terms == want["terms"], name
        assert rows == want["rows_sha256"], name
        assert coeffs == want["coeffs_sha256"], name
        assert stats == want["stats_sha256"], name
        assert ev == want["expectation_hex"], name
