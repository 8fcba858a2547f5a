"""Synthetic circuits of the propagate-path workloads (all random-free unless a seed is given).

* kicked-Ising chain: n qubits, open chain, per step ZZ rotations on every bond then X rotations on
  every qubit (``ExpGate`` = exp(+i theta H), each gate its own parameter);
* kicked-Ising square lattice: L x L open lattice, row-major sites, bonds (i, i+1) and (i, i+L) in
  row-major order of their first site;
* spinful Fermi-Hubbard on a periodic L x L lattice: mode 2*site+spin, Majoranas 2*mode and 2*mode+1;
  per step one hopping gate per bond per spin (generator i/2 (m_{2p} m_{2q+1} - m_{2p+1} m_{2q}) scaled
  to unit weight-2 terms) and one weight-4 on-site generator per site;
* brickwork ring: n qubits, per layer ZZ on even bonds, ZZ on odd bonds, X and Y on every qubit;
* parameter sweep over the kicked-Ising lattice: (theta_J, theta_h) sets uniform in [-pi/2, pi/2) from
  ``default_rng(2)``, eight observables (two each of Z_i, nearest-neighbour Z_i Z_j, plaquette ZZZZ
  and a horizontal X string of length 5) at sites from ``default_rng(3)``.
"""

from __future__ import annotations

import numpy as np

from .circuit import Circuit, ExpGate
from .majorana import MajoranaOperator
from .pauli import Pauli, PauliOperator


def _pauli_gate(letters: str, qubits, n: int) -> ExpGate:
    return ExpGate(PauliOperator({Pauli(letters, tuple(qubits)): 1.0}, n))


def chain_bonds(n: int) -> list[tuple[int, int]]:
    return [(i, i + 1) for i in range(n - 1)]


def lattice_bonds(L: int) -> list[tuple[int, int]]:
    bonds = []
    for r in range(L):
        for c in range(L):
            i = r * L + c
            if c + 1 < L:
                bonds.append((i, i + 1))
            if r + 1 < L:
                bonds.append((i, i + L))
    return bonds


def kicked_ising(n: int, bonds, steps: int, theta_j: float, theta_h: float) -> Circuit:
    gates, params = [], []
    for _ in range(steps):
        for i, j in bonds:
            gates.append(_pauli_gate("ZZ", (i, j), n))
            params.append(theta_j)
        for i in range(n):
            gates.append(_pauli_gate("X", (i,), n))
            params.append(theta_h)
    return Circuit(gates, n, params, initial_state=[])


def kicked_ising_chain(n: int = 20, steps: int = 5, theta_j: float = -np.pi / 4, theta_h: float = 0.3) -> Circuit:
    return kicked_ising(n, chain_bonds(n), steps, theta_j, theta_h)


def kicked_ising_lattice(L: int = 10, steps: int = 20, theta_j: float = -np.pi / 4, theta_h: float = 0.4) -> Circuit:
    return kicked_ising(L * L, lattice_bonds(L), steps, theta_j, theta_h)


def z_observable(n: int, sites, coeff: float = 1.0) -> PauliOperator:
    return PauliOperator({Pauli("Z", (s,)): coeff for s in sites}, n)


def mean_z(n: int) -> PauliOperator:
    return PauliOperator({Pauli("Z", (s,)): 1.0 / n for s in range(n)}, n)


def hubbard_lattice(L: int = 6, steps: int = 10, t_angle: float = 0.1, u_angle: float = 0.2):
    """Spinful Hubbard Trotter circuit on a periodic L x L lattice.

    Returns ``(circuit, n_modes)``.  Hopping gate per bond (p, q) and spin:
    ``MajoranaOperator({(2p, 2q+1): 1j/2... })`` written with unit-magnitude imaginary weight-2 terms
    ``{(2p, 2q+1): 1j, (2p+1, 2q): -1j}`` (two commuting Hermitian terms); on-site gate per site s:
    ``{(4s, 4s+1, 4s+2, 4s+3): 1.0}``.
    """
    sites = L * L
    n_modes = 2 * sites
    bonds = []
    for r in range(L):
        for c in range(L):
            s = r * L + c
            bonds.append((s, r * L + (c + 1) % L))
            bonds.append((s, ((r + 1) % L) * L + c))
    gates, params = [], []
    for _ in range(steps):
        for a, b in bonds:
            for spin in (0, 1):
                p, q = 2 * a + spin, 2 * b + spin
                terms = {(2 * p, 2 * q + 1): 1j, (2 * p + 1, 2 * q): -1j}
                gates.append(ExpGate(MajoranaOperator(terms, n_modes)))
                params.append(t_angle)
        for s in range(sites):
            gates.append(ExpGate(MajoranaOperator({(4 * s, 4 * s + 1, 4 * s + 2, 4 * s + 3): 1.0}, n_modes)))
            params.append(u_angle)
    occupied = [2 * s + (0 if (s // L + s % L) % 2 == 0 else 1) for s in range(sites)]
    return Circuit(gates, n_modes, params, initial_state=occupied), n_modes


def number_operator(mode: int, n_modes: int) -> MajoranaOperator:
    """n_j = 1/2 + (i/2) m_{2j} m_{2j+1} (ladder convention a_j = (m_{2j} + i m_{2j+1})/2)."""
    return MajoranaOperator({(): 0.5, (2 * mode, 2 * mode + 1): 0.5j}, n_modes)


def brickwork_ring(n: int = 64, layers: int = 12, seed: int = 0) -> Circuit:
    """Per layer: ZZ on bonds (2i, 2i+1), ZZ on bonds (2i+1, 2i+2 mod n), X and Y on every qubit;
    one independent angle per gate, uniform in [-pi, pi) from ``default_rng(seed)``."""
    gates = []
    for _ in range(layers):
        for parity in (0, 1):
            for i in range(parity, n, 2):
                j = (i + 1) % n
                gates.append(_pauli_gate("ZZ", (min(i, j), max(i, j)), n))
        for i in range(n):
            gates.append(_pauli_gate("X", (i,), n))
        for i in range(n):
            gates.append(_pauli_gate("Y", (i,), n))
    rng = np.random.default_rng(seed)
    params = rng.uniform(-np.pi, np.pi, len(gates))
    return Circuit(gates, n, params, initial_state=[])


def ring_zz_mean(n: int) -> PauliOperator:
    return PauliOperator({Pauli("ZZ", (min(i, (i + 1) % n), max(i, (i + 1) % n))): 1.0 / n for i in range(n)}, n)


def kicked_ising_rows(circuit: Circuit, sets) -> np.ndarray:
    """Per (theta_J, theta_h) set, the circuit's per-gate parameter row: theta_J on ZZ gates, theta_h on
    X gates (one row per set, ``circuit.n_parameters`` columns)."""
    is_zz = np.array([len(next(iter(g.generator.terms)).qubits) == 2 for g in circuit.gates])
    sets = np.asarray(sets, dtype=np.float64).reshape(-1, 2)
    return np.where(is_zz[None, :], sets[:, :1], sets[:, 1:])


def sweep_parameter_sets(n_sets: int = 256, seed: int = 2) -> np.ndarray:
    """``(n_sets, 2)`` array of (theta_J, theta_h), uniform in [-pi/2, pi/2)."""
    return np.random.default_rng(seed).uniform(-np.pi / 2, np.pi / 2, size=(n_sets, 2))


def sweep_observables(L: int = 10, seed: int = 3) -> list[PauliOperator]:
    """Eight observables of uneven weight on the L x L lattice (row-major sites): two single Z, two
    nearest-neighbour ZZ, two plaquettes ZZZZ, two horizontal X strings of length 5."""
    rng = np.random.default_rng(seed)
    n = L * L
    out = []
    for _ in range(2):
        out.append(PauliOperator({Pauli("Z", (int(rng.integers(n)),)): 1.0}, n))
    bonds = lattice_bonds(L)
    for _ in range(2):
        i, j = bonds[int(rng.integers(len(bonds)))]
        out.append(PauliOperator({Pauli("ZZ", (i, j)): 1.0}, n))
    for _ in range(2):
        r, c = int(rng.integers(L - 1)), int(rng.integers(L - 1))
        s = r * L + c
        out.append(PauliOperator({Pauli("ZZZZ", (s, s + 1, s + L, s + L + 1)): 1.0}, n))
    for _ in range(2):
        r, c = int(rng.integers(L)), int(rng.integers(L - 4))
        s = r * L + c
        out.append(PauliOperator({Pauli("XXXXX", tuple(range(s, s + 5))): 1.0}, n))
    return out

