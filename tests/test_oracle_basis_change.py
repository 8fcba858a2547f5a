"""The cutoff basis (basis_change; MonomialPropagator.inl:484-495, MonomialPropagatorCommon.h:44-63,
MajoranaAlgebra.h:178-189) in the oracle, pinned by

* the reference's own case (tests/test_basis.py:44-64): with the Jordan-Wigner basis the partner of
  one Majorana rotation is too long and is dropped, leaving cos(2) on the observable;
* the Jordan-Wigner equivalence (stated for the native Pauli encoding in
  pauli_build_layer_tests.cpp:325-390): a Majorana propagation whose support cutoff is measured in the
  JW basis keeps, gate by gate, exactly the terms that the native Pauli propagation of the JW-mapped
  circuit keeps (the Pauli arm is pinned by dense matrices and exhaustive algebra elsewhere), and both
  evolved operators are the same matrix.  Without the basis change the kept set differs, so the
  check is not vacuous.
"""

from __future__ import annotations

import numpy as np
import pytest

import dense
from monoprop_amd import Circuit, ExpGate, MajoranaOperator, Pauli, PauliOperator, jordan_wigner_basis_change
from monoprop_amd.circuit import expand_monomials


def test_reference_single_rotation_case(oracle_mod):
    n = 6
    o = oracle_mod.Oracle({(0,): 1.0}, n, basis="majorana", cutoff=n // 2, cutoff_type="length",
                          basis_change=jordan_wigner_basis_change(n))
    o.propagate([(5,)], [0], [-1.0], [1.0])
    ev = o.evolved_operator()
    assert list(ev) == [(0,)]
    assert np.isclose(ev[(0,)].real, np.cos(2 * 1.0))
    # the native Majorana length of the partner (0, 5) is 2 <= 3: kept without the basis change
    o2 = oracle_mod.Oracle({(0,): 1.0}, n, basis="majorana", cutoff=n // 2, cutoff_type="length")
    o2.propagate([(5,)], [0], [-1.0], [1.0])
    assert sorted(o2.evolved_operator()) == [(0,), (0, 5)]


def test_malformed_tables_rejected(oracle_mod):
    n = 3
    o = oracle_mod.Oracle({(0,): 1.0}, n, basis="majorana", cutoff=2, cutoff_type="length")
    for bad in ([[0]], [[0]] * (2 * n - 1), [[2 * n]] * (2 * n)):
        with pytest.raises(RuntimeError):
            o.set_basis_change(bad)
    p = oracle_mod.Oracle({(0, 1): 1.0}, 2, basis="pauli", cutoff=2)
    with pytest.raises(RuntimeError, match="Pauli"):
        p.set_basis_change(jordan_wigner_basis_change(2))


def _hermitian_coeff(idx, r):
    """Coefficient that makes r * (product over idx) Hermitian: M^dagger = (-1)^C(L,2) M."""
    L = len(idx)
    return r if (L * (L - 1) // 2) % 2 == 0 else 1j * r


def _as_pauli(idx, c, n, ms):
    """The JW image of c * m_idx as a single Pauli term {Pauli: real coefficient}."""
    terms = dense.pauli_decompose(c * dense.majorana_product(idx, ms), n)
    assert len(terms) == 1
    (slots, v), = terms.items()
    assert abs(v.imag) < 1e-12
    letters = dense.slots_to_letters(slots)
    qs = sorted(letters)
    return Pauli("".join(letters[q] for q in qs), tuple(qs)), float(v.real)


def _dense_majorana(ev, ms):
    return sum(v * dense.majorana_product(k, ms) for k, v in ev.items())


def _dense_pauli(ev, n):
    return sum(v * dense.pauli_matrix(dense.slots_to_letters(k), n) for k, v in ev.items())


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_jordan_wigner_cutoff_basis_equals_native_pauli(oracle_mod, seed):
    n, cutoff = 4, 3
    ms = dense.majorana_matrices(n)
    rng = np.random.default_rng(seed)
    maj_gates, pauli_gates, thetas = [], [], []
    for _ in range(12):
        L = int(rng.choice([1, 2, 2, 3, 4]))
        idx = tuple(sorted(int(i) for i in rng.choice(2 * n, size=L, replace=False)))
        c = _hermitian_coeff(idx, float(rng.uniform(0.5, 1.5)))
        maj_gates.append(ExpGate(MajoranaOperator({idx: c}, n)))
        p, v = _as_pauli(idx, c, n, ms)
        pauli_gates.append(ExpGate(PauliOperator({p: v}, n)))
        thetas.append(float(rng.uniform(-1.2, 1.2)))
    obs_idx, obs_c = (0, 3, 4, 6), 0.7
    p_obs, v_obs = _as_pauli(obs_idx, obs_c, n, ms)

    cm = Circuit(tuple(maj_gates), n)
    cp = Circuit(tuple(pauli_gates), n)
    mm, gm, pm, _ = expand_monomials(cm.gates, cm.resolved_mapping, n)
    mp_, gp, pp, _ = expand_monomials(cp.gates, cp.resolved_mapping, n)

    om = oracle_mod.Oracle({obs_idx: obs_c}, n, basis="majorana", cutoff=cutoff, cutoff_type="support",
                           basis_change=jordan_wigner_basis_change(n))
    op = oracle_mod.Oracle(PauliOperator({p_obs: v_obs}, n).get_local_operator().terms, n, basis="pauli",
                           cutoff=cutoff, cutoff_type="support")
    om.propagate(mm, pm, gm, thetas)
    op.propagate(mp_, pp, gp, thetas)
    for a, b in zip(om.gate_stats(), op.gate_stats()):
        assert np.array_equal(a, b)
    A = _dense_majorana(om.evolved_operator(0.0), ms)
    B = _dense_pauli(op.evolved_operator(0.0), n)
    assert np.allclose(A, B, atol=1e-11)

    # the truncation is active, and without the basis change it keeps a different set
    of = oracle_mod.Oracle({obs_idx: obs_c}, n, basis="majorana", cutoff=2 * n, cutoff_type="support",
                           basis_change=jordan_wigner_basis_change(n))
    of.propagate(mm, pm, gm, thetas)
    assert om.size() < of.size()
    on = oracle_mod.Oracle({obs_idx: obs_c}, n, basis="majorana", cutoff=cutoff, cutoff_type="support")
    on.propagate(mm, pm, gm, thetas)
    assert on.size() != om.size()

