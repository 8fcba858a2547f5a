"""Dense-matrix brute force used to pin the oracle (small qubit/mode counts only).

Pauli strings and Jordan-Wigner Majoranas are built as explicit 2^n x 2^n matrices; a circuit is
applied in the Heisenberg picture as ``O -> U^dag O U`` with ``U = exp(+i theta g P)``.  Nothing here
follows the engine's bit layout or sign tables, so it is an independent check of them.
"""

from __future__ import annotations

import itertools
from functools import reduce

import numpy as np

I2 = np.eye(2, dtype=complex)
X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)
LETTER = {"I": I2, "X": X, "Y": Y, "Z": Z}


def kron_all(mats):
    return reduce(np.kron, mats)


def pauli_matrix(letters: dict[int, str], n: int) -> np.ndarray:
    """Qubit 0 is the leftmost tensor factor."""
    return kron_all([LETTER[letters.get(q, "I")] for q in range(n)])


def slots_to_letters(slots) -> dict[int, str]:
    present = set(slots)
    out = {}
    for q in {s // 2 for s in present}:
        u, v = (2 * q) in present, (2 * q + 1) in present
        out[q] = "Z" if (u and v) else ("X" if u else "Y")
    return out


def letters_to_slots(letters: dict[int, str]) -> tuple[int, ...]:
    out = []
    for q, l in letters.items():
        if l == "X":
            out.append(2 * q)
        elif l == "Y":
            out.append(2 * q + 1)
        elif l == "Z":
            out.extend((2 * q, 2 * q + 1))
    return tuple(sorted(out))


def majorana_matrices(n: int) -> list[np.ndarray]:
    """m_{2q} = Z..Z X_q, m_{2q+1} = Z..Z Y_q (Jordan-Wigner)."""
    ms = []
    for q in range(n):
        for L in ("X", "Y"):
            ms.append(kron_all([Z] * q + [LETTER[L]] + [I2] * (n - q - 1)))
    return ms


def majorana_product(idx, ms) -> np.ndarray:
    out = np.eye(ms[0].shape[0], dtype=complex)
    for i in idx:
        out = out @ ms[i]
    return out


def expm_herm(h: np.ndarray, theta: float) -> np.ndarray:
    """exp(i theta h) for Hermitian h."""
    w, v = np.linalg.eigh(h)
    return (v * np.exp(1j * theta * w)) @ v.conj().T


def heisenberg(op: np.ndarray, gens: list[np.ndarray], thetas: list[float]) -> np.ndarray:
    """U^dag O U for U = U_L ... U_1, U_k = exp(i theta_k H_k) (circuit order)."""
    u = np.eye(op.shape[0], dtype=complex)
    for h, t in zip(gens, thetas):
        u = expm_herm(h, t) @ u
    return u.conj().T @ op @ u


def pauli_decompose(m: np.ndarray, n: int, tol: float = 1e-13) -> dict[tuple[int, ...], complex]:
    """Slot-tuple -> coefficient of the Pauli expansion of m."""
    out = {}
    for combo in itertools.product("IXYZ", repeat=n):
        letters = {q: l for q, l in enumerate(combo) if l != "I"}
        p = pauli_matrix(letters, n)
        c = np.trace(p @ m) / (2**n)
        if abs(c) > tol:
            out[letters_to_slots(letters)] = c
    return out


def majorana_decompose(m: np.ndarray, n: int, tol: float = 1e-13) -> dict[tuple[int, ...], complex]:
    """Sorted Majorana index tuple -> coefficient (basis m_S are orthogonal under Tr(A^dag B)/2^n)."""
    ms = majorana_matrices(n)
    out = {}
    for w in range(2 * n + 1):
        for idx in itertools.combinations(range(2 * n), w):
            b = majorana_product(idx, ms)
            c = np.trace(b.conj().T @ m) / (2**n)
            if abs(c) > tol:
                out[idx] = c
    return out


def statevector_expectation(n: int, obs_letters: list[tuple[dict[int, str], float]], gates, thetas,
                            occupied=()) -> float:
    """<b|U^dag O U|b> by statevector for product-basis |b> (qubit q set iff q in occupied).

    ``gates`` is a list of ``letters`` dicts (single Pauli-string generators with g = 1)."""
    psi = np.zeros(2**n, dtype=complex)
    idx = 0
    for q in occupied:
        idx |= 1 << (n - 1 - q)
    psi[idx] = 1.0
    for letters, t in zip(gates, thetas):
        psi = apply_pauli_rotation(psi, letters, t, n)
    val = 0.0
    for letters, c in obs_letters:
        val += c * np.vdot(psi, apply_pauli(psi, letters, n)).real
    return float(val)


def apply_pauli(psi: np.ndarray, letters: dict[int, str], n: int) -> np.ndarray:
    out = psi.reshape([2] * n)
    for q, l in letters.items():
        out = np.moveaxis(np.tensordot(LETTER[l], out, axes=([1], [q])), 0, q)
    return out.reshape(-1)


def apply_pauli_rotation(psi, letters, theta, n):
    """exp(i theta P) psi = cos(theta) psi + i sin(theta) P psi."""
    return np.cos(theta) * psi + 1j * np.sin(theta) * apply_pauli(psi, letters, n)
