"""Loader for the reference's msgpack problem fixtures (schema: reference tests/data/README.md).

The files under ``tests/golden/reference_fixtures/`` are the reference's own test data, copied
unchanged: exact energies and gradients of small fermionic circuits at infinite cutoff.
"""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import msgpack
import numpy as np

from monoprop_amd.circuit import Circuit
from monoprop_amd.majorana import MajoranaOperator

DATA = Path(__file__).resolve().parent / "golden" / "reference_fixtures"
EXACT_CASES = ["lih_fermionic_spin_exact", "rx_rz_ry_rz_exact", "random_exact"]


@dataclass
class Problem:
    name: str
    majoranas: list
    gen_coeffs: np.ndarray
    param_inds: np.ndarray
    parameters: np.ndarray
    initial_state: list
    n_modes: int
    operator: MajoranaOperator
    exact_energy: float
    exact_gradient: np.ndarray

    def circuit(self) -> Circuit:
        return Circuit.from_dense_arrays(self.majoranas, self.gen_coeffs, self.param_inds, self.n_modes,
                                         self.parameters, self.initial_state)


def load(name: str) -> Problem:
    data = msgpack.unpackb((DATA / f"{name}.msgpack").read_bytes())
    ham = data["hamiltonian"]
    terms = {tuple(k): complex(r, i) for k, r, i in zip(ham["keys"], ham["real"], ham["imag"], strict=True)}
    n = int(data["num_modes"])
    return Problem(
        name=name,
        majoranas=[tuple(m) for m in data["majoranas"]],
        gen_coeffs=np.asarray(data["gen_coeffs"], dtype=np.float64),
        param_inds=np.asarray(data["param_inds"], dtype=np.int64),
        parameters=np.asarray(data["parameters"], dtype=np.float64),
        initial_state=list(data["hartree_fock"]),
        n_modes=n,
        operator=MajoranaOperator(terms, n),
        exact_energy=float(data["actual_energy"]),
        exact_gradient=np.asarray(data["actual_gradient"], dtype=np.float64),
    )
