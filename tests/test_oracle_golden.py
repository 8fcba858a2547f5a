"""The oracle against its committed golden vectors (tests/golden/oracle_vectors.json, written by
tests/golden/make_golden.py): same term count, same store and coefficients to the bit (SHA-256 of
the arrays in store order), same <O> bit for bit, and for the small cases the same evolved operator."""

from __future__ import annotations

import importlib.util
import json
import os
from pathlib import Path

import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
_spec = importlib.util.spec_from_file_location("make_golden", GOLDEN / "make_golden.py")
make_golden = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(make_golden)
VECTORS = json.loads((GOLDEN / "oracle_vectors.json").read_text())


@pytest.mark.parametrize("name", sorted(VECTORS))
def test_oracle_reproduces_golden(oracle_mod, name):
    case = make_golden.cases()[name]
    if case.get("heavy") and os.environ.get("MPA_GOLDEN_HEAVY") != "1":
        pytest.skip("heavy workload case (tens of seconds of oracle time); MPA_GOLDEN_HEAVY=1 runs it")
    o = make_golden.run_oracle(case)
    got = make_golden.record(o, keep_terms="evolved" in VECTORS[name])
    assert got == VECTORS[name]


def test_every_case_is_recorded():
    assert set(make_golden.cases()) == set(VECTORS)
