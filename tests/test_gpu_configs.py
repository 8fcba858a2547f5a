"""The benchmark configurations at their own sizes, on the HIP engine.

* 10x10 kicked-Ising lattice, all 20 Trotter steps, theta_h 0.1 / 0.4 / 0.8, the three as members of
  one batched store (the timed workload): every member's rows and FP64 coefficients hash to the
  oracle's golden digests and its <O> equals the oracle's to the bit (the lone propagations are held
  to the same digests by test_gpu_golden.py);
* the parameter sweep's own inputs (its eight observables at its first two parameter sets, cutoff 8,
  lower_atol 1e-5, 3 Trotter steps): one batch of 16 members, the sweep driver forced to split the
  batch on the device, and two drivers at once, against the oracle's golden digests and <O> bits;
* 64-qubit brickwork ring, all 12 layers (about 2300 angles), the gradient of the graph built at the
  seeded angles: <O> of the replay against the oracle's propagation (1e-10 relative: the replay
  applies the rotations without the fused sweep's (v k)(1/k) recovery), the adjoint gradient against
  the frozen structure's parameter-shift rule on 32 sampled angles (each angle drives one layer that is
  linear in cos 2a and sin 2a, so dE/dtheta = E(theta + pi/4) - E(theta - pi/4) exactly; 1e-12) and
  against central differences with h = 1e-5 (1e-8).
"""

from __future__ import annotations

import importlib.util
import json
from pathlib import Path

import numpy as np
import pytest

from monoprop_amd import PauliPropagator, workloads
from monoprop_amd.batch import BatchPropagator, expectation_grid

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
_spec = importlib.util.spec_from_file_location("make_golden", GOLDEN / "make_golden.py")
make_golden = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(make_golden)
VECTORS = json.loads((GOLDEN / "oracle_vectors.json").read_text())


def _member_matches(w, c, m, q, want):
    sel = m == q
    assert int(sel.sum()) == want["terms"], (q, int(sel.sum()), want["terms"])
    assert make_golden.digest(np.ascontiguousarray(w[sel], dtype=np.uint64)) == want["rows_sha256"], q
    assert make_golden.digest(np.ascontiguousarray(c[sel], dtype=np.float64)) == want["coeffs_sha256"], q


def test_lattice10_full_depth_batched():
    print("case: lattice 10x10, 20 steps, theta_h 0.1/0.4/0.8 batched", flush=True)
    circ = workloads.kicked_ising_lattice(10, 20)
    names = ["lattice10_full_h01", "lattice10_full_h04", "lattice10_full_h08"]
    sets = [(-np.pi / 4, 0.1), (-np.pi / 4, 0.4), (-np.pi / 4, 0.8)]
    b = BatchPropagator([workloads.mean_z(100)] * 3, [], cutoff=10, lower_atol=1e-5)
    b.propagate(circ, workloads.kicked_ising_rows(circ, sets))
    w, c, m = b._sim.rows()
    ev = b.expectation_values()
    for q, name in enumerate(names):
        _member_matches(w, c, m, q, VECTORS[name])
        assert float(ev[q]).hex() == VECTORS[name]["expectation_hex"], (name, ev[q])
    # the parameter-sweep row builder gives the lone circuits' own parameters
    for th, row in zip((0.1, 0.4, 0.8), workloads.kicked_ising_rows(circ, sets)):
        assert np.array_equal(row, workloads.kicked_ising_lattice(10, 20, theta_h=th).parameters)


def _sweep_inputs():
    circ = workloads.kicked_ising_lattice(10, 3)
    obs = workloads.sweep_observables()
    sets = workloads.sweep_parameter_sets()[:2]
    return circ, obs, sets


def test_sweep_inputs_one_batch():
    print("case: sweep observables x 2 parameter sets, 3 steps, one batch of 16", flush=True)
    circ, obs, sets = _sweep_inputs()
    ops = [o for _ in range(2) for o in obs]
    params = workloads.kicked_ising_rows(circ, [sets[i] for i in range(2) for _ in obs])
    b = BatchPropagator(ops, [], cutoff=8, lower_atol=1e-5)
    b.propagate(circ, params)
    w, c, m = b._sim.rows()
    ev = b.expectation_values()
    for q in range(len(ops)):
        name = f"sweep_set{q // len(obs)}_obs{q % len(obs)}"
        _member_matches(w, c, m, q, VECTORS[name])
        assert float(ev[q]).hex() == VECTORS[name]["expectation_hex"], (name, ev[q])


def test_sweep_inputs_with_device_splits():
    """The sweep driver with a small capacity bound: batches stop, split on the device and resume;
    every <O> keeps the oracle's bits."""
    print("case: sweep driver, 2 sets x 8 observables, forced splits", flush=True)
    circ, obs, sets = _sweep_inputs()
    stats = []
    grid = expectation_grid(circ, obs, workloads.kicked_ising_rows(circ, sets), cutoff=8, lower_atol=1e-5,
                            max_members=16, max_capacity=1 << 21, stats=stats)
    assert any(s[3] for s in stats), "the bound should have forced at least one split"
    for i in range(2):
        for j in range(len(obs)):
            assert float(grid[i, j]).hex() == VECTORS[f"sweep_set{i}_obs{j}"]["expectation_hex"], (i, j)


@pytest.mark.parametrize("max_capacity", [-1, 1 << 21])
def test_sweep_inputs_concurrent_drivers(max_capacity):
    """Two sweep drivers at once (workers=2, each set on its own engines and stream; the automatic bound
    shares the memory, the small one forces splits): every <O> keeps the oracle's bits."""
    print(f"case: sweep driver, 2 workers, max_capacity={max_capacity}", flush=True)
    circ, obs, sets = _sweep_inputs()
    grid = expectation_grid(circ, obs, workloads.kicked_ising_rows(circ, sets), cutoff=8, lower_atol=1e-5,
                            max_members=16, max_capacity=max_capacity, workers=2)
    for i in range(2):
        for j in range(len(obs)):
            assert float(grid[i, j]).hex() == VECTORS[f"sweep_set{i}_obs{j}"]["expectation_hex"], (i, j)


def test_ring64_full_gradient():
    print("case: ring 64, 12 layers, full gradient", flush=True)
    circ = workloads.brickwork_ring(64, 12, 0)
    prop = PauliPropagator(workloads.ring_zz_mean(64), [], cutoff=8, lower_atol=1e-6)
    prop.build_graph(circ)
    theta = np.asarray(circ.parameters, dtype=np.float64)
    assert len(theta) > 2000
    ref = float.fromhex(VECTORS["ring64_12layers"]["expectation_hex"])
    val, grad = prop.expectation_value_and_gradient(theta)
    assert abs(val - ref) <= 1e-10 * abs(ref), (val, ref)
    assert val == prop.expectation_value(theta)
    ks = np.random.default_rng(1).choice(len(theta), 32, replace=False)
    for k in ks:
        up, dn = theta.copy(), theta.copy()
        up[k] += np.pi / 4
        dn[k] -= np.pi / 4
        shift = prop.expectation_value(up) - prop.expectation_value(dn)
        assert abs(grad[k] - shift) <= 1e-12, (k, grad[k], shift)
        h = 1e-5
        up, dn = theta.copy(), theta.copy()
        up[k] += h
        dn[k] -= h
        fd = (prop.expectation_value(up) - prop.expectation_value(dn)) / (2 * h)
        assert abs(grad[k] - fd) <= 1e-8, (k, grad[k], fd)
