"""The HIP engine against the committed golden vectors of the oracle (no oracle at run time): the
term store, the FP64 coefficients and the per-gate counters hash to the recorded SHA-256 digests,
and <O> matches to the bit (both sum the terms in ascending row order, as the reference does).

The vectors include the benchmark configurations at full depth (the 10x10 kicked-Ising lattice over
20 Trotter steps at every theta_h, the 64-qubit ring over 12 layers) and at the depth the oracle
covers (the 6x6 Hubbard lattice, 3 Trotter steps; the sweep's observables and first parameter sets,
3 Trotter steps)."""

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


@pytest.mark.parametrize("name", sorted(VECTORS))
def test_engine_matches_golden(name):
    print(f"case: golden {name}", flush=True)
    case = make_golden.cases()[name]
    want = VECTORS[name]
    circ = make_golden.build_circuit(case["circuit"])
    monos, gcs, pm, _ = make_golden.layers(circ)
    d = DeviceSimulator(case["obs"], case["cutoff"], case["state"], lower_atol=case["lower"],
                        upper_atol=case["upper"], cutoff_type=case["cutoff_type"], logical_num_modes=case["n"],
                        basis=case["basis"], schrodinger_cutoff=case["schr"])
    d.propagate(monos, pm, gcs, make_golden.parameters(case, circ))
    w, c = d.rows()
    assert d.size() == want["terms"]
    assert make_golden.digest(np.ascontiguousarray(w, dtype=np.uint64)) == want["rows_sha256"]
    assert make_golden.digest(np.ascontiguousarray(c, dtype=np.float64)) == want["coeffs_sha256"]
    n, a, e, i = d.gate_stats()
    assert int(np.sum(n)) == want["term_gates"]
    assert make_golden.stats_digest(n, a, e, i) == want["stats_sha256"]
    ev = d.expectation_value([])
    assert ev.hex() == want["expectation_hex"], (ev, float.fromhex(want["expectation_hex"]))
    if "evolved" in want:
        assert d.evolved_operator([], 0.0) == {tuple(k): complex(*v) for k, v in want["evolved"]}
