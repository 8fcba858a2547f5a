"""The allocation-free <O> entry point (mpa_expectation_value_noalloc) through the C ABI: equal to the
bit to mpa_expectation_value on two calls in a row, the second on a smaller operator (the static
scratch array keeps the first call's contents), on a lone Pauli and a lone Majorana propagator; its
error paths (stored graph, Schrodinger picture, too many terms are refused)."""

from __future__ import annotations

import ctypes as C

import pytest

from monoprop_amd import workloads
from monoprop_amd._engine import DeviceSimulator, load_library
from monoprop_amd.circuit import expand_monomials

pytestmark = pytest.mark.gpu


def _both(d):
    lib = load_library()
    a = C.c_double()
    assert lib.mpa_expectation_value(d._h, None, 0, C.byref(a)) == 0, lib.mpa_last_error()
    b = C.c_double()
    assert lib.mpa_expectation_value_noalloc(d._h, C.byref(b)) == 0, lib.mpa_last_error()
    return a.value, b.value


def _lattice(L, steps, th):
    circ = workloads.kicked_ising_lattice(L, steps, theta_h=th)
    monos, gcs, pm, _ = expand_monomials(circ.gates, circ.resolved_mapping, L * L)
    d = DeviceSimulator(workloads.mean_z(L * L).get_local_operator().terms, 10, [], lower_atol=1e-5,
                        cutoff_type="support", logical_num_modes=L * L, basis="pauli")
    d.propagate(monos, pm, gcs, circ.parameters)
    return d


def test_noalloc_matches_expectation_twice_smaller_second():
    print("case: noalloc <O>, 10x10 lattice 6 steps then 4x4 lattice 3 steps", flush=True)
    big = _lattice(10, 6, 0.4)
    small = _lattice(4, 3, 0.4)
    assert small.size() < big.size()
    for d in (big, small):
        ref, got = _both(d)
        assert got.hex() == ref.hex(), (got, ref)


def test_noalloc_majorana_and_state_phases():
    print("case: noalloc <O>, 3x3 Hubbard 2 steps", flush=True)
    circ, n = workloads.hubbard_lattice(3, 2)
    monos, gcs, pm, _ = expand_monomials(circ.gates, circ.resolved_mapping, n)
    d = DeviceSimulator(workloads.number_operator(2, n).terms, 6, list(circ.initial_state), lower_atol=1e-6,
                        cutoff_type="length", logical_num_modes=n, basis="majorana")
    d.propagate(monos, pm, gcs, circ.parameters)
    ref, got = _both(d)
    assert got.hex() == ref.hex(), (got, ref)


def test_noalloc_refuses_what_it_cannot_serve():
    print("case: noalloc <O> errors", flush=True)
    lib = load_library()
    out = C.c_double()
    d = DeviceSimulator({(0,): 1.0}, 4, [], logical_num_modes=4, basis="pauli", cutoff_type="support")
    d.build_graph([(2,)], [0], [1.0])
    assert lib.mpa_expectation_value_noalloc(d._h, C.byref(out)) != 0
    assert b"contracted operator" in lib.mpa_last_error()
    s = DeviceSimulator({(0,): 1.0}, 4, [], logical_num_modes=4, basis="pauli", cutoff_type="support",
                        schrodinger_cutoff=2)
    assert lib.mpa_expectation_value_noalloc(s._h, C.byref(out)) != 0
    assert b"Heisenberg" in lib.mpa_last_error()


def test_noalloc_refuses_more_terms_than_its_array():
    """2^20 + 1 fully paired initial terms (Z strings on 21 qubits): one more than the array holds."""
    print("case: noalloc <O> with 2^20 + 1 terms", flush=True)
    n = 21
    terms = {}
    for i in range(1, (1 << 20) + 2):
        terms[tuple(s for q in range(n) if (i >> q) & 1 for s in (2 * q, 2 * q + 1))] = 1.0
    d = DeviceSimulator(terms, n, [], logical_num_modes=n, basis="pauli", cutoff_type="support")
    lib = load_library()
    out = C.c_double()
    assert lib.mpa_expectation_value_noalloc(d._h, C.byref(out)) != 0
    assert b"more than the 1048576" in lib.mpa_last_error()
    ref = C.c_double()
    assert lib.mpa_expectation_value(d._h, None, 0, C.byref(ref)) == 0
    assert ref.value == float(len(terms))
