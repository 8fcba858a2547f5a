"""The basis-change helpers of the reference's utils module (src/monoprop/utils.py:43-72), with the
cases of its tests/test_utils.py:64-120, and the Jordan-Wigner table checked against dense matrices."""

from __future__ import annotations

import numpy as np
import pytest

import dense
from monoprop_amd.utils import jordan_wigner_basis_change, validate_basis_change


def test_single_qubit_has_no_z_prefix():
    assert jordan_wigner_basis_change(1) == [[0], [1]]


def test_multi_qubit_has_growing_z_string():
    basis = jordan_wigner_basis_change(3)
    assert basis == [[0], [1], [0, 1, 2], [0, 1, 3], [0, 1, 2, 3, 4], [0, 1, 2, 3, 5]]


def test_zero_qubits_is_empty():
    assert jordan_wigner_basis_change(0) == []


def test_validate_accepts_none_and_matching_length():
    assert validate_basis_change(None, 4) is None
    assert validate_basis_change(None, 0) is None
    assert validate_basis_change(jordan_wigner_basis_change(3), 3) is None


def test_validate_rejects_wrong_length():
    with pytest.raises(ValueError, match=r"Basis change must have length 6, but got 4\."):
        validate_basis_change(jordan_wigner_basis_change(2), 3)
    with pytest.raises(ValueError, match=r"Basis change must have length 2, but got 0\."):
        validate_basis_change([], 1)


@pytest.mark.parametrize("n", [1, 2, 3])
def test_rows_are_the_jordan_wigner_images(n):
    """Row k is the slot support (X_q -> 2q, Y_q -> 2q+1, Z_q -> both) of the Pauli string that the
    dense Jordan-Wigner Majorana m_k is, up to a phase."""
    ms = dense.majorana_matrices(n)
    for k, row in enumerate(jordan_wigner_basis_change(n)):
        terms = dense.pauli_decompose(ms[k], n)
        assert len(terms) == 1
        (slots, v), = terms.items()
        assert abs(abs(v) - 1.0) < 1e-12
        assert sorted(row) == sorted(slots)
        assert np.allclose(ms[k], v * dense.pauli_matrix(dense.slots_to_letters(row), n))
