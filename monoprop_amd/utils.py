"""Small shared validators (mirror of reference ``src/monoprop/utils.py``)."""

from __future__ import annotations


def validate_system_size(size: int, *, argument_name: str) -> int:
    """Return ``size`` as int; reject bools, non-integers and negatives."""
    if isinstance(size, bool) or not isinstance(size, int):
        raise TypeError(f"{argument_name} must be an integer (not {type(size).__name__}).")
    if size < 0:
        raise ValueError(f"{argument_name} must be non-negative; got {size}.")
    return int(size)


def validate_basis_change(basis_change, num_modes: int) -> None:
    """Accept None or a table of exactly ``2 * num_modes`` rows (reference ``utils.py:60-72``); the
    engine checks the slots themselves when the table is set."""
    if basis_change is not None and len(basis_change) != 2 * num_modes:
        raise ValueError(f"Basis change must have length {2 * num_modes}, but got {len(basis_change)}.")


def is_antihermitian(indices) -> bool:
    """Whether the Majorana product over ``indices`` is anti-Hermitian: (m_a1 ... m_aL)^dagger equals
    (-1)^C(L,2) times the product, so it is anti-Hermitian iff C(L,2) = L(L-1)/2 is odd, i.e. iff L/2
    (integer division) is odd (reference ``MajoranaAlgebra.h:44-47``, exported from ``bindings.cpp.in:123-126``)."""
    return (len(indices) // 2) % 2 == 1


def antihermitian_generator_correction(indices) -> complex:
    """``c = i^((C(L,2) + 1) mod 4)`` for a Majorana product M of ``L = len(indices)`` distinct indices:
    since M^dagger = (-1)^C(L,2) M, (c M)^dagger = -c M, so c M is anti-Hermitian and exp(theta c M) is
    unitary for real theta (reference ``MajoranaAlgebra.h:49-51``, exported from
    ``bindings.cpp.in:127-130``)."""
    n = len(indices)
    return (1, 1j, -1, -1j)[(n * (n - 1) // 2 + 1) % 4]


def jordan_wigner_basis_change(n_qubits: int) -> list[list[int]]:
    """Slot support of the Jordan-Wigner image of each Majorana ``m_k`` in the native Pauli slots."""
    out: list[list[int]] = []
    for i in range(n_qubits):
        prefix = list(range(2 * i))
        out.append(prefix + [2 * i])
        out.append(prefix + [2 * i + 1])
    return out
