"""The reference's hand-computed evolution of a non-fermionic Majorana operator
(tests/test_nonfermi.py:21-62) as a golden vector: the oracle (CPU) and the HIP engine (GPU) must both
reproduce it to 1e-12, with the same number of terms."""

from __future__ import annotations

import numpy as np
import pytest

from monoprop_amd import Circuit, MajoranaOperator, MajoranaPropagator
from monoprop_amd.circuit import expand_monomials

N_MODES = 2
CIRCUIT = dict(majoranas=[(0,), (0, 1), (0, 1, 3), (2, 3)], gen_coeffs=np.array([-1.0, 1.0, -1.0, 1.0]),
               param_inds=np.array([0, 1, 2, 3]), system_size=N_MODES, parameters=np.array([0.1, 0.2, 0.2, 0.2]),
               initial_state=[])
OBSERVABLE = {(1, 2, 3): -1j}
EXACT = {
    (1, 2, 3): -0.8314427691150754j,
    (0, 1, 2, 3): -0.16854179325074592 + 0j,
    (0, 2, 3): 0.35867804544976145j,
    (0, 2): 0.35152836455073294j,
    (2,): 0.07125832726038464 + 0j,
    (1, 2): 0.1516466453264173j,
}


def _check(evolved: dict):
    assert len(evolved) == len(EXACT)
    for k, v in EXACT.items():
        assert abs(complex(evolved[k]) - v) <= 1e-12, k


def test_oracle_reproduces_the_hand_computed_evolution(oracle_mod):
    c = Circuit.from_dense_arrays(**CIRCUIT)
    monos, gcs, pm, _ = expand_monomials(c.gates, c.resolved_mapping, N_MODES)
    o = oracle_mod.Oracle(MajoranaOperator(OBSERVABLE, N_MODES).terms, N_MODES, basis="majorana", cutoff=4,
                          cutoff_type="length")
    o.propagate(monos, pm, gcs, c.parameters)
    _check({k: v for k, v in o.evolved_operator().items() if k})


@pytest.mark.gpu
def test_engine_reproduces_the_hand_computed_evolution():
    print("case: nonfermi", flush=True)
    c = Circuit.from_dense_arrays(**CIRCUIT)
    mp = MajoranaPropagator(MajoranaOperator(OBSERVABLE, N_MODES), c.initial_state, cutoff=4)
    mp.propagate(c)
    ev = mp.evolved_operator()
    _check(ev.terms)
    assert ev.isclose(MajoranaOperator(EXACT, N_MODES), atol=1e-12)
