"""Benchmark: 100-qubit 2D kicked-Ising Pauli propagation on one MI355X.

Workload (one "step"): the 10x10 open-lattice kicked-Ising circuit (180 ZZ bonds, row-major), 20
Trotter steps of ZZ rotations theta_J = -pi/4 then X rotations theta_h, propagated once for each
theta_h in {0.1, 0.4, 0.8}; observable (1/100) sum_i Z_i, support cutoff 10, lower_atol 1e-5,
FP64 coefficients, two u64 limbs per plane.  Inputs are synthetic circuits (no data files).  By
default the three propagations run at once, each on its own engine (own HIP stream, own host thread);
`schedules` times the same steps as the three members of one batched store (each member's terms and
coefficients are those of its own propagation) and as one propagation after another.

Metric: term-gate applications per second, T / t with T = sum over gates of the live terms before
the gate (counted on the device); t covers engine creation, propagation and <O> for the three
circuits (their generator lists are built once, before timing).  The roofline figure is one gate's two
kernels (sweep and resolve) together: their algorithmic bytes over their HIP-event-timed device
duration, from one extra step of the same
workload run with the per-launch profiler on (kept out of the timed steps; its propagations run one
after another, so that each launch's duration is its own).  Algorithmic bytes per gate
are the path's measurement definition: the sweep reads 16 B of x/z limbs per live term (per limb pair
that meets the generator), reads and writes each anticommuting coefficient, reads the anticommuting
terms' other limbs and writes each emitted partner record (row + value); the resolve reads those records
back and appends each new term.  `hbm_GBps` is what the kernel really moves (PMC counters).
`per_propagation` lists the engines of the last timed step (theta_h, term-gate applications, device
time of the gate loop): the step lasts at least as long as the largest.  `gradient` reports the
other half of the metric, the full-angle gradient's time against one forward replay, on the 64-qubit
ring brickwork (2304 angles), measured after the timed steps; `batch` times a parameter sweep (sets x
eight observables, 10 Trotter steps) batched against one propagation per pair.  The CPU baseline is the oracle
(a restatement of the reference algorithm) on the timed workload itself, run on the host's thread
budget, its stores and <O> compared bit for bit with the GPU step's (the serial rate on a sample beside it).

`--schrodinger Q` measures the Schrodinger picture instead (state terms of at most Q qubits to start,
one propagation at a time, default theta_h 0.8: at 0.1 and 0.4 the 100-qubit state outgrows the GPU).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--trotter-steps 20] [--no-cpu-baseline]
       [--no-gradient] [--batch-sets 4] [--batched | --unbatched] [--schrodinger Q] [--theta-h 0.1 0.4 0.8]
       [--only-gradient] [--only-batch]
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
THETA_H = (0.1, 0.4, 0.8)
SCHRODINGER_QUBITS: int | None = None  # --schrodinger: the picture and its state cutoff in qubits


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--trotter-steps", type=int, default=20)
    ap.add_argument("--lattice", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-gradient", action="store_true")
    ap.add_argument("--batch-sets", type=int, default=4, help="parameter sets of the batched sweep (0: skip)")
    ap.add_argument("--batch-no-solo", action="store_true", help="skip the one-propagation-per-pair comparison")
    ap.add_argument("--batch-workers", type=int, default=8, help="sweep drivers running at once")
    ap.add_argument("--only-batch", action="store_true", help="run and print only the batched sweep")
    ap.add_argument("--only-gradient", action="store_true", help="run and print only the gradient measurement")
    ap.add_argument("--copy-calibration", action="store_true", help="only run the known-size copy (counter scale)")
    ap.add_argument("--traffic-from", nargs=4, metavar=("COPY_FETCH", "COPY_WRITE", "FETCH", "WRITE"),
                    help="derive per-launch HBM traffic from rocprofv3 --pmc CSVs and write the summary")
    ap.add_argument("--traffic-out", default=str(TRAFFIC_FILE))
    ap.add_argument("--cpu-sample-steps", type=int, default=4, help="Trotter steps of the serial CPU sample")
    ap.add_argument("--batched", action="store_true",
                    help="time the step as one batched store (the theta_h as its members)")
    ap.add_argument("--unbatched", action="store_true",
                    help="time the step as one propagation per theta_h, one after another (default: the three "
                         "at once, each on its own engine and stream)")
    ap.add_argument("--no-compare", action="store_true", help="skip timing the other step schedules for comparison")
    ap.add_argument("--schrodinger", type=int, default=None, metavar="Q",
                    help="Schrodinger picture, initial state terms of at most Q qubits (one propagation at a time)")
    ap.add_argument("--theta-h", type=float, nargs="+", default=None, help="X angles of the step's propagations")
    return ap.parse_args()


def layers_of(circ):
    from monoprop_amd.circuit import expand_monomials

    return expand_monomials(circ.gates, circ.resolved_mapping, circ.system_size)


def observable_terms(n):
    from monoprop_amd import workloads

    return workloads.mean_z(n).get_local_operator().terms


def build_inputs(L, trotter):
    """The three circuits' generator lists (host-side input, built once outside the timed region)."""
    from monoprop_amd import workloads

    out = []
    for th in THETA_H:
        circ = workloads.kicked_ising_lattice(L, trotter, theta_h=th)
        monos, gcs, pm, _ = layers_of(circ)
        out.append((monos, gcs, pm, list(circ.parameters)))
    return out


def run_gpu_step_batched(L, inputs, profile):
    """One step as one store: the three theta_h propagations are members of one batch (same gates,
    their own angles), so each gate is one pair of launches for all three.  Returns (term-gate
    applications, <O> list, [simulator])."""
    from monoprop_amd._engine import BatchSimulator

    n = L * L
    monos, gcs, pm, _ = inputs[0]
    params = np.asarray([p for _, _, _, p in inputs], dtype=np.float64)
    sim = BatchSimulator([observable_terms(n)] * len(inputs), 10, [], lower_atol=1e-5, cutoff_type="support",
                         logical_num_modes=n, basis="pauli")
    if profile:
        sim.set_profiling(True)
    sim.propagate(monos, pm, gcs, params)
    total_t = int(sim.gate_stats()[0].sum())
    return total_t, [float(v) for v in sim.expectation_values()], [sim]


def _lone_propagation(n, inp, profile):
    """One theta_h propagated by its own engine: (term-gate applications, <O>, simulator)."""
    from monoprop_amd._engine import DeviceSimulator

    monos, gcs, pm, params = inp
    schr = None if SCHRODINGER_QUBITS is None else 2 * SCHRODINGER_QUBITS
    sim = DeviceSimulator(observable_terms(n), 10, [], lower_atol=1e-5, cutoff_type="support",
                          logical_num_modes=n, basis="pauli", schrodinger_cutoff=schr)
    if profile:
        sim.set_profiling(True)
    sim.propagate(monos, pm, gcs, params)
    return int(sim.gate_stats()[0].sum()), sim.expectation_value([]), sim


def run_gpu_step(L, inputs, profile):
    """One step: a propagation per theta_h, one after another; returns (term-gate applications, <O> list,
    simulators)."""
    res = [_lone_propagation(L * L, inp, profile) for inp in inputs]
    return sum(r[0] for r in res), [r[1] for r in res], [r[2] for r in res]


def run_gpu_step_concurrent(L, inputs, profile):
    """One step: a propagation per theta_h, each by its own engine (own HIP stream) on its own host
    thread, all at once (monoprop_amd.concurrent); returns (term-gate applications, <O> list, simulators).
    Profiled steps run one after another, so that each launch's event-timed duration is its own."""
    if profile:
        return run_gpu_step(L, inputs, profile)
    from monoprop_amd.concurrent import run_concurrently

    res = run_concurrently([lambda inp=inp: _lone_propagation(L * L, inp, False) for inp in inputs])
    return sum(r[0] for r in res), [r[1] for r in res], [r[2] for r in res]


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _host_threads() -> int:
    """Threads for the CPU baseline: the job's thread budget (OMP_NUM_THREADS), else the cores."""
    try:
        return max(1, int(os.environ.get("OMP_NUM_THREADS", "")))
    except ValueError:
        return max(1, os.cpu_count() or 1)


def _digests(words, coeffs) -> tuple[str, str]:
    import hashlib

    return (hashlib.sha256(np.ascontiguousarray(words, dtype=np.uint64).tobytes()).hexdigest(),
            hashlib.sha256(np.ascontiguousarray(coeffs, dtype=np.float64).tobytes()).hexdigest())


def gpu_member_results(L, inputs, step):
    """One step of the timed workload; per theta_h its store digests and <O> bits (a batched step's
    members are cut out of the shared store)."""
    _, evs, sims = step(L, inputs, profile=False)
    out = []
    if step is run_gpu_step_batched:
        w, c, m = sims[0].rows()
        for q, ev in enumerate(evs):
            sel = m == q
            out.append((*_digests(w[sel], c[sel]), float(ev).hex()))
    else:
        for sim, ev in zip(sims, evs):
            out.append((*_digests(*sim.rows()), float(ev).hex()))
    return out


def cpu_baseline(L, trotter, sample_steps, gpu_results=None):
    """The oracle on the timed workload itself (every theta_h, all Trotter steps) on the host's thread
    budget, its stores and <O> compared with the GPU step's; the serial oracle (1 core) on a sample
    beside it: theta_h = 0.4, the first `sample_steps` Trotter steps.  The Schrodinger mode times only
    its sample."""
    import oracle
    from monoprop_amd import workloads

    n = L * L
    schr = None if SCHRODINGER_QUBITS is None else 2 * SCHRODINGER_QUBITS
    picture = "" if schr is None else f", Schrodinger picture (state terms of <= {SCHRODINGER_QUBITS} qubits)"

    def run(th, steps, threads):
        circ = workloads.kicked_ising_lattice(L, steps, theta_h=th)
        monos, gcs, pm, _ = layers_of(circ)
        oracle.set_threads(threads)
        try:
            o = oracle.Oracle(observable_terms(n), n, basis="pauli", cutoff=10, cutoff_type="support",
                              lower_atol=1e-5, schrodinger_cutoff=schr)
            t0 = time.perf_counter()
            o.propagate(monos, pm, gcs, circ.parameters)
            dt = time.perf_counter() - t0
        finally:
            oracle.set_threads(1)
        return int(o.gate_stats()[0].sum()), dt, o

    threads = _host_threads()
    th_sample = 0.4 if SCHRODINGER_QUBITS is None else THETA_H[0]
    if SCHRODINGER_QUBITS is None:
        total_t, total_s, same = 0, 0.0, True
        for q, th in enumerate(THETA_H):
            t, dt, o = run(th, trotter, threads)
            total_t += t
            total_s += dt
            if gpu_results is not None:
                mine = (*_digests(*o.rows()), float(o.expectation()).hex())
                same = same and mine == gpu_results[q]
            del o
        value = total_t / total_s
        sample = (f"the timed workload: {L}x{L} kicked Ising, {trotter} Trotter steps, theta_h "
                  f"{{{','.join(str(t) for t in THETA_H)}}} ({total_t} term-gate applications, {total_s:.1f} s), "
                  f"oracle restatement on {threads} threads")
    else:
        t, dt, _ = run(th_sample, sample_steps, threads)
        value, same = t / dt, None
        sample = (f"{L}x{L} kicked Ising theta_h={th_sample}, first {sample_steps} of {trotter} Trotter steps "
                  f"({t} term-gate applications, {dt:.1f} s){picture}, oracle restatement on {threads} threads")
    st, sdt, _ = run(th_sample, sample_steps, 1)
    out = {
        "value": value,
        "unit": "term-gate applications/s",
        "cores": threads,
        "kind": "port",
        "sample": sample,
        "cpu_model": _cpu_model(),
        "serial": {"value": st / sdt, "cores": 1,
                   "sample": f"theta_h={th_sample}, first {sample_steps} of {trotter} Trotter steps ({st} "
                             f"term-gate applications, {sdt:.1f} s){picture}, serial oracle"},
    }
    if same is not None and gpu_results is not None:
        # every member's rows, coefficients (SHA-256) and <O> bits equal to the GPU step's
        out["matches_gpu"] = bool(same)
    return out


def gradient_measurement(reps: int = 3):
    """64-qubit ring brickwork, 12 layers (2304 independent angles), O = mean ZZ, cutoff 8, lower_atol
    1e-6: build the graph once, then time the replay (<O>) against replay + adjoint (<O>, d<O>/dtheta)."""
    from monoprop_amd import workloads
    from monoprop_amd._engine import DeviceSimulator

    circ = workloads.brickwork_ring(64, 12, seed=0)
    monos, gcs, pm, gidx = layers_of(circ)
    theta = np.asarray(circ.parameters, dtype=np.float64)
    sim = DeviceSimulator(workloads.ring_zz_mean(64).get_local_operator().terms, 8, [], lower_atol=1e-6,
                          cutoff_type="support", logical_num_modes=64, basis="pauli")
    t0 = time.perf_counter()
    sim.build_graph(monos, pm, gcs, gidx, list(theta))
    build_s = time.perf_counter() - t0
    fwd, grd = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        sim.expectation_value(theta)
        fwd.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        sim.expectation_value_and_gradient(theta)
        grd.append(time.perf_counter() - t0)
    f, g = min(fwd), min(grd)
    cos_records, rot_records = sim.graph_size()
    segs = sim.gradient_segments()
    # the same gradient with the least adjoint memory (more segments, each earlier one replayed again)
    sim.set_gradient_checkpointing(False, 1)
    low = []
    for _ in range(reps):
        t0 = time.perf_counter()
        _, g_low = sim.expectation_value_and_gradient(theta)
        low.append(time.perf_counter() - t0)
    low_segs = sim.gradient_segments()
    return {
        "workload": "64-qubit ring brickwork, 12 layers, 2304 angles, O=mean ZZ, cutoff 8, lower_atol 1e-6",
        "terms": sim.size(),
        "graph_records": {"cos": cos_records, "rotation": rot_records},
        "graph_build_ms": build_s * 1e3,
        "forward_ms": f * 1e3,
        "gradient_ms": g * 1e3,
        "gradient_over_forward": g / f,
        # device memory of the adjoint beyond the coefficient vector: the graph (records) and, per
        # gradient, the checkpoints and one segment's pre-layer values
        "memory_GB": {"graph": sim.graph_memory_bytes() / 1e9, "checkpoints": segs["checkpoint_bytes"] / 1e9,
                      "snapshots": segs["snapshot_bytes"] / 1e9, "coefficient_vector": sim.size() * 8 / 1e9},
        "segments": segs["segments"],
        "least_memory": {"segments": low_segs["segments"], "gradient_ms": min(low) * 1e3,
                         "gradient_over_forward": min(low) / f,
                         "checkpoints_GB": low_segs["checkpoint_bytes"] / 1e9,
                         "snapshots_GB": low_segs["snapshot_bytes"] / 1e9},
    }


def batch_measurement(n_sets: int, solo: bool = True, workers: int = 4):
    """Parameter sweep over the 10x10 kicked-Ising lattice, 10 Trotter steps, cutoff 8, lower_atol 1e-5:
    the first `n_sets` (theta_J, theta_h) sets of default_rng(2) x the eight seeded observables, once
    through the batched sweep driver (`workers` drivers at once) and once as one propagation per pair."""
    from monoprop_amd import workloads
    from monoprop_amd._engine import DeviceSimulator
    from monoprop_amd.batch import expectation_grid

    n = 100
    circ = workloads.kicked_ising_lattice(10, 10)
    obs = workloads.sweep_observables()
    rows = workloads.kicked_ising_rows(circ, workloads.sweep_parameter_sets()[:n_sets])
    stats = []
    t_start = time.perf_counter()

    def progress(e):  # one line per propagate call, so a long sweep keeps reporting
        print(f"[{time.perf_counter() - t_start:7.1f} s] batch of {e[0]} members: {e[1]} terms, {e[2]:.0f} ms"
              f"{', split' if e[3] else ''}", file=sys.stderr, flush=True)

    grid = expectation_grid(circ, obs, rows, cutoff=8, lower_atol=1e-5, stats=stats, progress=progress,
                            workers=workers)
    t_batch = time.perf_counter() - t_start
    tg = sum(s[4] for s in stats)
    out = {
        "workload": f"10x10 kicked-Ising, 10 Trotter steps, {n_sets} (theta_J, theta_h) sets x 8 observables, "
                    "cutoff 8, lower_atol 1e-5",
        "drivers": workers,
        "propagations": int(grid.size),
        "batched_s": t_batch,
        "batched_term_gates_per_s": tg / t_batch,
        "batched_device_s": sum(s[2] for s in stats) / 1e3,  # summed over the drivers' propagate calls
        "propagate_calls": len(stats),
        "splits": sum(1 for s in stats if s[3]),
        "term_gate_applications": tg,
    }
    if not solo:
        return out
    monos, gcs, pm, _ = layers_of(circ)
    ref, t_solo, tg_solo = np.empty_like(grid), 0.0, 0
    for i in range(n_sets):
        for j, o in enumerate(obs):
            t0 = time.perf_counter()
            d = DeviceSimulator(o.get_local_operator().terms, 8, [], lower_atol=1e-5, cutoff_type="support",
                                logical_num_modes=n, basis="pauli")
            d.propagate(monos, pm, gcs, list(rows[i]))
            ref[i, j] = d.expectation_value([])
            t_solo += time.perf_counter() - t0
            tg_solo += int(d.gate_stats()[0].sum())
            del d
    out.update(solo_s=t_solo, solo_term_gates_per_s=tg_solo / t_solo,
               max_abs_diff_vs_solo=float(np.max(np.abs(grid - ref))))
    return out


COPY_BYTES = 256 << 20
TRAFFIC_FILE = ROOT / "profiles" / "traffic_kicked_ising_10x10.json"


def copy_calibration():
    """A plain device copy of COPY_BYTES, three times: the known-size reference for the HBM counters."""
    import torch

    a = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {COPY_BYTES} bytes x3")


def traffic_from_counters(copy_fetch, copy_write, fetch, write, out_path=TRAFFIC_FILE):
    """Per-launch HBM bytes of the fused kernels from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes
    (KiB per dispatch) over this benchmark, each scaled by what the counter reads for the copy above."""
    import csv
    from collections import defaultdict

    def per_kernel(path):
        out = defaultdict(list)
        for r in csv.DictReader(open(path)):
            for k in ("k_sweep", "k_resolve", "copyBuffer"):
                if k in r["Kernel_Name"]:
                    out[k].append(float(r["Counter_Value"]) * 1024.0)
        return out

    cf, cw = per_kernel(copy_fetch)["copyBuffer"], per_kernel(copy_write)["copyBuffer"]
    out = {"fetch_scale": COPY_BYTES / (sum(cf) / len(cf)), "write_scale": COPY_BYTES / (sum(cw) / len(cw)),
           "kernels": {}}
    fk, wk = per_kernel(fetch), per_kernel(write)
    for k in ("k_sweep", "k_resolve"):
        fv = [x for x in fk[k] if x > 0]  # a stopped device leaves a few empty launches behind
        wv = [x for x in wk[k] if x > 0]
        fb = out["fetch_scale"] * sum(fv) / max(len(fv), 1)
        wb = out["write_scale"] * sum(wv) / max(len(wv), 1)
        out["kernels"][k] = {"fetch_bytes_per_launch": fb, "write_bytes_per_launch": wb,
                             "traffic_bytes_per_launch": fb + wb, "launches": len(fv)}
    Path(out_path).write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out, indent=1))


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary, or None."""
    try:
        return json.loads(TRAFFIC_FILE.read_text())["kernels"][kernel]["traffic_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def propagate_merge_fractions(prof):
    """The propagate (k_sweep) and merge (k_resolve) kernels together: algorithmic bytes and PMC
    traffic over their summed device time, each as a fraction of the HBM peak."""
    ks = [k for k in ("k_sweep", "k_resolve") if k in prof]
    ms = sum(prof[k]["ms"] for k in ks)
    if not ks or ms <= 0:
        return None
    alg = sum(prof[k]["bytes"] for k in ks) / (ms * 1e-3) / 1e9
    out = {"ms": round(ms, 3), "algorithmic_GBps": round(alg, 1), "algorithmic_frac": round(alg / HBM_PEAK_GBS, 4)}
    traffic = [measured_traffic(k) for k in ks] if SCHRODINGER_QUBITS is None else [None]
    if all(t is not None for t in traffic):
        hbm = sum(t * prof[k]["launches"] for t, k in zip(traffic, ks)) / (ms * 1e-3) / 1e9
        out.update({"hbm_GBps": round(hbm, 1), "hbm_frac": round(hbm / HBM_PEAK_GBS, 4)})
    return out


def step_rate(prof, ms_per_step):
    """Device-wide rate of one timed step: the sweep and resolve launches' algorithmic bytes and PMC
    traffic per step over the wall time per step."""
    ks = [k for k in ("k_sweep", "k_resolve") if k in prof]
    if not ks or ms_per_step <= 0:
        return None
    alg = sum(prof[k]["bytes"] for k in ks) / (ms_per_step * 1e-3) / 1e9
    out = {"algorithmic_GBps": round(alg, 1), "algorithmic_frac": round(alg / HBM_PEAK_GBS, 4)}
    traffic = [measured_traffic(k) for k in ks] if SCHRODINGER_QUBITS is None else [None]
    if all(t is not None for t in traffic):
        hbm = sum(t * prof[k]["launches"] for t, k in zip(traffic, ks)) / (ms_per_step * 1e-3) / 1e9
        out.update({"hbm_GBps": round(hbm, 1), "hbm_frac": round(hbm / HBM_PEAK_GBS, 4)})
    return out


def main():
    global THETA_H, SCHRODINGER_QUBITS
    args = parse_args()
    if args.schrodinger is not None:
        SCHRODINGER_QUBITS = args.schrodinger
        THETA_H = (0.8,)
        args.unbatched, args.batched = True, False  # one propagation: the batched store is Heisenberg only
    if args.theta_h:
        THETA_H = tuple(args.theta_h)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.traffic_from:
        traffic_from_counters(*args.traffic_from, out_path=args.traffic_out)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the engine holds the whole operator in "
              "one GPU's HBM and multi-GPU sharding is not part of this build.", file=sys.stderr)
        sys.exit(2)
    import torch

    from monoprop_amd import _engine

    _engine.load_library()
    if args.only_gradient:
        print(json.dumps({"gradient": gradient_measurement()}))
        return
    if args.only_batch:
        print(json.dumps({"batch": batch_measurement(args.batch_sets, solo=not args.batch_no_solo,
                                                     workers=args.batch_workers)}))
        return
    inputs = build_inputs(args.lattice, args.trotter_steps)
    schedules = {"concurrent": run_gpu_step_concurrent, "batched": run_gpu_step_batched,
                 "sequential": run_gpu_step}
    mode = "sequential" if args.unbatched else "batched" if args.batched else "concurrent"
    step = schedules[mode]

    def timed(step_fn):
        torch.cuda.synchronize()
        for _ in range(args.warmup):
            step_fn(args.lattice, inputs, profile=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        total, evs = 0, None
        for i in range(args.steps):
            t, evs, sims = step_fn(args.lattice, inputs, profile=False)
            total += t
            if i == args.steps - 1:
                last[:] = sims  # the last step's engines, read after the timed region
            del sims
        torch.cuda.synchronize()
        return total, evs, time.perf_counter() - t0

    last: list = []
    total_t, evs, elapsed = timed(step)
    # each engine of the last timed step: its term-gate applications and device time (HIP events around
    # its gate loop)
    per_propagation = [{"term_gate_applications": int(sim.gate_stats()[0].sum()),
                        "device_ms": round(sim.last_propagate_ms(), 2)} for sim in last]
    if len(per_propagation) == len(THETA_H):  # one engine per theta_h (not the batched store)
        for rec, th in zip(per_propagation, THETA_H):
            rec["theta_h"] = th
    last.clear()

    # per-kernel device time (HIP events around every launch) and algorithmic bytes, from one more step
    # of the same workload run with the profiler on (it is kept out of the timed steps above)
    prof: dict[str, dict[str, float]] = {}
    _, _, sims = step(args.lattice, inputs, profile=True)
    prof_wall_ms = 0.0
    for sim in sims:
        prof_wall_ms += sim.last_propagate_ms()
        for k, v in sim.kernel_profile().items():
            acc = prof.setdefault(k, {"ms": 0.0, "bytes": 0.0, "launches": 0})
            for f in acc:
                acc[f] += v[f]
    del sims
    prof = {k: v for k, v in prof.items() if v["launches"] > 0}

    # The roofline unit: one gate of the fused pipeline, its propagate (k_sweep) and merge (k_resolve)
    # launches together.  The path's algorithmic bytes are defined per gate over both, and the two share
    # the device time about evenly; without them (other pipelines) the kernel with the most device time.
    # The PMC traffic file covers the default (Heisenberg) workload only.
    if "k_sweep" in prof and "k_resolve" in prof:
        dname = "k_sweep + k_resolve"
        parts = [prof["k_sweep"], prof["k_resolve"]]
        d = {"ms": sum(x["ms"] for x in parts), "bytes": sum(x["bytes"] for x in parts),
             "launches": prof["k_sweep"]["launches"]}
        tr = [measured_traffic(k) for k in ("k_sweep", "k_resolve")]
        traffic = sum(tr) if SCHRODINGER_QUBITS is None and all(t is not None for t in tr) else None
    else:
        dname, d = max(prof.items(), key=lambda kv: kv[1]["ms"])
        traffic = measured_traffic(dname) if SCHRODINGER_QUBITS is None else None
    avg_ms = d["ms"] / max(d["launches"], 1)
    achieved = d["bytes"] / (d["ms"] * 1e-3) / 1e9 if d["ms"] > 0 else 0.0
    kern_ms = sum(v["ms"] for v in prof.values())
    kern_bytes = sum(v["bytes"] for v in prof.values())

    line = {
        "metric": "term-gate applications/s",
        "value": total_t / elapsed,
        "unit": "term-gate applications/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.lattice}x{args.lattice} kicked-Ising Pauli propagation, {args.trotter_steps} Trotter "
                        f"steps, theta_J=-pi/4, theta_h in {{{','.join(str(t) for t in THETA_H)}}}, O=mean Z, "
                        "cutoff 10, lower_atol 1e-5"
                        + ("" if SCHRODINGER_QUBITS is None else
                           f", Schrodinger picture (initial state terms of <= {SCHRODINGER_QUBITS} qubits)"),
            "qubits": args.lattice * args.lattice,
            "gates_per_propagation": (2 * args.lattice * (args.lattice - 1) + args.lattice ** 2) * args.trotter_steps,
            "term_gate_applications_per_step": total_t // args.steps,
            "expectation": evs,
            "schedule": mode,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dname,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "avg_launch_ms": avg_ms,  # per gate: the sweep's and the resolve's launch together
            "launches": d["launches"],
            # what the kernels actually move through HBM (PMC traffic over the event-timed launches); the
            # column index needs far fewer bytes than the limb reads of the algorithmic figure
            "hbm_GBps": (traffic / (avg_ms * 1e-3) / 1e9) if traffic else None,
            "hbm_frac": (traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else None,
        },
        "kernels": {
            k: {"ms": round(v["ms"], 3), "GB/s": round(v["bytes"] / (v["ms"] * 1e6), 1) if v["ms"] > 0 else 0.0,
                "launches": v["launches"],
                "hbm_GBps": (round(measured_traffic(k) / (v["ms"] / v["launches"] * 1e-3) / 1e9, 1)
                             if SCHRODINGER_QUBITS is None and measured_traffic(k) and v["launches"] else None)}
            for k, v in sorted(prof.items(), key=lambda kv: -kv[1]["ms"])
        },
        "kernel_ms_per_step": round(kern_ms, 3),
        # the engines of the last timed step: the step lasts at least as long as the largest of them
        "per_propagation": per_propagation,
        "propagate_merge": propagate_merge_fractions(prof),
        # the whole device over the timed step: one step's algorithmic bytes and PMC traffic of the sweep
        # and resolve launches (from the profiled step: the same gates) over the timed wall time per step,
        # which with the concurrent schedule covers three propagations overlapping on the device
        "step_rate": step_rate(prof, elapsed * 1e3 / args.steps),
        "all_kernels_GBps": kern_bytes / (kern_ms * 1e6) if kern_ms > 0 else 0.0,
        "profiled_step": {"propagate_ms": round(prof_wall_ms, 2), "kernel_ms": round(kern_ms, 2)},
    }
    if not args.no_compare and SCHRODINGER_QUBITS is None:  # the same steps under the other schedules
        line["schedules"] = {}
        for other, fn in schedules.items():
            if other == mode:
                continue
            t_u, evs_u, el_u = timed(fn)
            last.clear()
            line["schedules"][other] = {"value": t_u / el_u, "ms_per_step": el_u * 1e3 / args.steps,
                                        "same_expectation_bits": [float(a).hex() == float(b).hex()
                                                                  for a, b in zip(evs, evs_u)]}
    if not args.no_gradient:
        line["gradient"] = gradient_measurement()
    if args.batch_sets > 0:
        line["batch"] = batch_measurement(args.batch_sets, solo=not args.batch_no_solo, workers=args.batch_workers)
    if not args.no_cpu_baseline:
        gpu_res = None if SCHRODINGER_QUBITS is not None else gpu_member_results(args.lattice, inputs, step)
        line["cpu_baseline"] = cpu_baseline(args.lattice, args.trotter_steps, args.cpu_sample_steps, gpu_res)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
