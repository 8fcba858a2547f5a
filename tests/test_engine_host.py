"""Host-side conversion of a propagate call's gate list (no GPU): the index tuples become CSR offsets
and int32 indices, the parameter mapping int64 and the generator coefficients float64, exactly as the
engine's C ABI (`mpa_propagate`) reads them."""

from __future__ import annotations

import numpy as np
import pytest

from monoprop_amd import workloads
from monoprop_amd._engine import DeviceSimulator, _csr
from monoprop_amd.circuit import expand_monomials


def test_csr_small_cases():
    offs, flat = _csr([(0, 1, 2), (), (5,)])
    assert offs.dtype == np.int64 and flat.dtype == np.int32
    assert offs.tolist() == [0, 3, 3, 4]
    assert flat.tolist() == [0, 1, 2, 5]
    # numpy integers and arrays are accepted as indices and tuples
    offs, flat = _csr([np.array([7, 3], dtype=np.int16), (np.int64(9),)])
    assert offs.tolist() == [0, 2, 3] and flat.tolist() == [7, 3, 9]


@pytest.mark.parametrize("tuples", [[], [()], [(), ()]])
def test_csr_without_indices(tuples):
    offs, flat = _csr(tuples)
    assert offs.tolist() == [0] * (len(tuples) + 1)
    assert flat.tolist() == [0] and flat.dtype == np.int32  # a valid pointer for the C call


def test_csr_rejects_negative_index():
    with pytest.raises(ValueError, match=r"index -2 is out of range; must be non-negative"):
        _csr([(1, 3), (4, -2)])


def test_layers_of_the_kicked_ising_circuit():
    """The 10x10 kicked-Ising gate list (5600 monomials): offsets are the running sums of the tuple
    lengths and the flat indices the tuples concatenated in order."""
    circ = workloads.kicked_ising_lattice(10, 20, theta_h=0.4)
    monos, gcs, pm, _ = expand_monomials(circ.gates, circ.resolved_mapping, circ.system_size)
    offs, slots, pmv, gc = DeviceSimulator._layers(monos, pm, gcs)
    assert len(offs) == len(monos) + 1 and offs[0] == 0
    for k in (0, 1, 2, len(monos) // 2, len(monos) - 1):
        assert slots[offs[k]:offs[k + 1]].tolist() == [int(i) for i in monos[k]]
    assert np.all(np.diff(offs) == np.array([len(m) for m in monos]))
    assert pmv.dtype == np.int64 and pmv.tolist() == [int(p) for p in pm]
    assert gc.dtype == np.float64 and gc.tolist() == [float(c) for c in gcs]


def test_layers_length_mismatch():
    with pytest.raises(ValueError, match="one entry per monomial"):
        DeviceSimulator._layers([(0,), (1,)], [0], [1.0, 1.0])
