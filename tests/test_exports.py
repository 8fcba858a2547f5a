"""Module attributes of the reference package (src/monoprop/__init__.py:21-70) that sit on the propagate
path, and the two Majorana helpers of its compiled core (bindings.cpp.in:123-130), checked against
dense Jordan-Wigner matrices."""

from __future__ import annotations

import itertools

import numpy as np
import pytest

import monoprop_amd as mp

# the reference's __all__ without the fermionic / integral / qiskit adapters, which are not on the path
REFERENCE_NAMES = [
    "MAX_NUM_MODES", "Circuit", "ExpGate", "Majorana", "MajoranaOperator", "MajoranaPropagator",
    "MonomialPropagator", "Pauli", "PauliOperator", "PauliPropagator", "__build_type__", "__compiler_flags__",
    "__variant__", "__version__", "antihermitian_generator_correction", "expand_monomials", "has_mpi",
    "is_antihermitian", "jordan_wigner_basis_change", "validate_parameter_mapping",
]


def test_reference_names_exported():
    for name in REFERENCE_NAMES:
        assert hasattr(mp, name), name
        assert name in mp.__all__, name
    assert mp.has_mpi is False
    from monoprop_amd._engine import DeviceSimulator

    assert mp.MAX_NUM_MODES == 250  # monoprop_MAX_NUM_MODES (reference CMakeLists.txt:45-50)
    assert DeviceSimulator.storage_num_modes == 256
    assert "--offload-arch=gfx950" in mp.__compiler_flags__


def _majoranas(n_qubits: int):
    """m_{2k} = Z...Z X_k, m_{2k+1} = Z...Z Y_k (dense, qubit 0 leftmost)."""
    I = np.eye(2)
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    Y = np.array([[0, -1j], [1j, 0]])
    Z = np.diag([1.0, -1.0]).astype(complex)

    def kron(ops):
        out = np.array([[1.0 + 0j]])
        for o in ops:
            out = np.kron(out, o)
        return out

    ms = []
    for k in range(n_qubits):
        for P in (X, Y):
            ms.append(kron([Z] * k + [P] + [I] * (n_qubits - k - 1)))
    return ms


@pytest.mark.parametrize("n_qubits", [3])
def test_antihermitian_helpers_against_dense_products(n_qubits):
    ms = _majoranas(n_qubits)
    dim = 2 ** n_qubits
    for L in range(0, 2 * n_qubits + 1):
        for idx in itertools.combinations(range(2 * n_qubits), L):
            M = np.eye(dim, dtype=complex)
            for i in idx:
                M = M @ ms[i]
            anti = np.allclose(M.conj().T, -M)
            herm = np.allclose(M.conj().T, M)
            assert anti != herm
            assert mp.is_antihermitian(idx) == anti, idx
            c = mp.antihermitian_generator_correction(idx)
            cM = c * M
            assert np.allclose(cM.conj().T, -cM), (idx, c)
            assert abs(abs(c) - 1.0) == 0.0


def test_dispatch_mode_count_errors():
    """The reference's dispatch checks (tools/generate-dispatch.py:145-160): a ValueError subclass
    with its messages, for a non-positive count and for more than MAX_NUM_MODES modes."""
    from monoprop_amd._engine import dispatch
    from monoprop_amd.exceptions import NumberOfModesInvalidError

    assert issubclass(NumberOfModesInvalidError, ValueError)
    for bad in (0, -3):
        with pytest.raises(NumberOfModesInvalidError, match=f"Number of Fermionic modes {bad} invalid. num_modes must be > 0."):
            dispatch(bad)
    with pytest.raises(NumberOfModesInvalidError, match="Number of Fermionic modes 251 invalid. num_modes must be <= 250. "
                                                       "Contact monoprop developers if more than 250 Fermionic modes "
                                                       "are required."):
        dispatch(251)
    with pytest.raises(NumberOfModesInvalidError):
        dispatch(2.0)
    assert callable(dispatch(250))
    with pytest.raises(NumberOfModesInvalidError):
        mp.PauliPropagator(mp.PauliOperator({"Z": 1.0}, 300), [], cutoff=4)
