"""Encodings between Pauli letters, native Pauli slots and Jordan-Wigner Majorana indices.

The engine's native Pauli encoding gives every qubit two slots (reference
``src/monoprop/conversion_utils.py:66-114``): ``X_q -> (2q,)``, ``Y_q -> (2q+1,)``,
``Z_q -> (2q, 2q+1)``.  On the device a Pauli string is stored as two bit planes (x, z); the slot
tuple is only the boundary format.
"""

from __future__ import annotations

from collections.abc import Sequence

_LETTER_SLOTS = {"X": (0,), "Y": (1,), "Z": (0, 1)}


def pauli_to_slots(string: str, qubits: Sequence[int]) -> tuple[int, ...]:
    """Native slot tuple of a local Pauli term (sorted)."""
    if len(string) != len(qubits):
        raise ValueError("Pauli string and qubits must have the same length")
    out: list[int] = []
    for letter, q in zip(string, qubits):
        for off in _LETTER_SLOTS.get(letter, ()):
            out.append(2 * int(q) + off)
    return tuple(sorted(out))


def slots_to_pauli(slots: Sequence[int]) -> tuple[str, tuple[int, ...]]:
    """Inverse of :func:`pauli_to_slots`: ``(letters, qubits)`` sorted by qubit."""
    present = set(int(s) for s in slots)
    letters = []
    qubits = []
    for q in sorted({s // 2 for s in present}):
        u, v = (2 * q) in present, (2 * q + 1) in present
        letters.append("Z" if (u and v) else ("X" if u else "Y"))
        qubits.append(q)
    return "".join(letters), tuple(qubits)


def permutation_sign(seq: Sequence[int]) -> int:
    """(-1)^(number of inversions) of a sequence."""
    inv = 0
    n = len(seq)
    for i in range(n):
        for j in range(i + 1, n):
            if seq[i] > seq[j]:
                inv += 1
    return -1 if inv % 2 else 1


def cancel_pairs(sorted_idx: Sequence[int]) -> tuple[int, ...]:
    """Drop adjacent equal indices in pairs (``m_i m_i = 1``)."""
    out: list[int] = []
    for i in sorted_idx:
        if out and out[-1] == i:
            out.pop()
        else:
            out.append(i)
    return tuple(out)


def pauli_string_to_majorana(full: str) -> tuple[tuple[int, ...], complex]:
    """Jordan-Wigner image of a full-width Pauli string as ``(Majorana indices, phase)``.

    With ``m_{2q} = Z_0..Z_{q-1} X_q`` and ``m_{2q+1} = Z_0..Z_{q-1} Y_q``: ``Z_q = -i m_{2q} m_{2q+1}``,
    ``X_q = (prod_{p<q} Z_p) m_{2q}``, ``Y_q = (prod_{p<q} Z_p) m_{2q+1}``.  The product is assembled
    letter by letter and canonicalized with its reordering sign.
    """
    idx: list[int] = []
    phase: complex = 1.0 + 0j
    for q, letter in enumerate(full):
        if letter == "I":
            continue
        if letter == "Z":
            factor = [2 * q, 2 * q + 1]
            phase *= -1j
        else:
            zs: list[int] = []
            for p in range(q):
                zs.extend((2 * p, 2 * p + 1))
            phase *= (-1j) ** q
            factor = zs + [2 * q + (0 if letter == "X" else 1)]
        idx.extend(factor)
    sign = permutation_sign(idx)
    return cancel_pairs(sorted(idx)), phase * sign
