"""Small shared validators (mirror of reference ``src/monoprop/utils.py``)."""

from __future__ import annotations


def validate_system_size(size: int, *, argument_name: str) -> int:
    """Return ``size`` as int; reject bools, non-integers and negatives."""
    if isinstance(size, bool) or not isinstance(size, int):
        raise TypeError(f"{argument_name} must be an integer (not {type(size).__name__}).")
    if size < 0:
        raise ValueError(f"{argument_name} must be non-negative; got {size}.")
    return int(size)


def jordan_wigner_basis_change(n_qubits: int) -> list[list[int]]:
    """Slot support of the Jordan-Wigner image of each Majorana ``m_k`` in the native Pauli slots."""
    out: list[list[int]] = []
    for i in range(n_qubits):
        prefix = list(range(2 * i))
        out.append(prefix + [2 * i])
        out.append(prefix + [2 * i + 1])
    return out
