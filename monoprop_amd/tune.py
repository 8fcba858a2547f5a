"""Autotuner of the fused pipeline's launch parameters.

Times the 10x10 kicked-Ising step (the three theta_h propagations, cutoff 10, lower_atol 1e-5, run at
once on their own engines and streams, the benchmark's schedule) for every pair of

* ``resolve_grid``: k_resolve blocks that take emitter chunks, and
* ``min_group_log2``: the smallest count group (2^g sweep tiles of 1024 terms), which sets how many
  group prefixes every resolve block scans against whether an emitter's sweep block follows from the
  prefixes alone (one-block groups, g = 2) or from a round of sweep-block sums (larger groups),

checks that every pair gives the same <O> to the bit (the results do not depend on launch
parameters), and writes the table and the fastest pair to ``tuning.json`` next to the library, which
every engine created afterwards applies (``_engine.default_launch``).

Usage: python -m monoprop_amd.tune [--trotter-steps 20] [--repeats 2] [--out monoprop_amd/tuning.json]
"""

from __future__ import annotations

import argparse
import itertools
import json
import sys
import time
from pathlib import Path

from . import _engine, workloads
from .circuit import expand_monomials
from .concurrent import run_concurrently

RESOLVE_GRIDS = (64, 96, 128, 192, 256, 320)
MIN_GROUP_LOG2 = (2, 3, 4, 5)
DEFAULT = (128, 2)  # the engine's built-in launch parameters
THETA_H = (0.1, 0.4, 0.8)


def _inputs(L: int, steps: int):
    monos = gcs = pm = None
    rows = []
    for th in THETA_H:
        circ = workloads.kicked_ising_lattice(L, steps, theta_h=th)
        monos, gcs, pm, _ = expand_monomials(circ.gates, circ.resolved_mapping, circ.system_size)
        rows.append(list(circ.parameters))
    return monos, gcs, pm, rows


def _one(n, monos, gcs, pm, params, launch):
    obs = workloads.mean_z(n).get_local_operator().terms
    sim = _engine.DeviceSimulator(obs, 10, [], lower_atol=1e-5, cutoff_type="support", logical_num_modes=n,
                                  basis="pauli")
    sim.set_launch(*launch)
    sim.propagate(monos, pm, gcs, params)
    return sim.expectation_value([])


def _step(L, inputs, launch):
    monos, gcs, pm, rows = inputs
    t0 = time.perf_counter()
    ev = run_concurrently([lambda p=p: _one(L * L, monos, gcs, pm, p, launch) for p in rows])
    return time.perf_counter() - t0, [float(v) for v in ev]


def tune(L: int = 10, steps: int = 20, repeats: int = 2, progress=None) -> dict:
    _engine.load_library()
    inputs = _inputs(L, steps)
    table, ref = [], None
    _step(L, inputs, DEFAULT)  # warm the device memory cache
    for grid, g in itertools.product(RESOLVE_GRIDS, MIN_GROUP_LOG2):
        times = []
        for _ in range(repeats):
            dt, ev = _step(L, inputs, (grid, g))
            times.append(dt)
            if ref is None:
                ref = ev
            elif ev != ref:
                raise RuntimeError(f"launch ({grid}, {g}) changed <O>: {ev} vs {ref}")
        row = {"resolve_grid": grid, "min_group_log2": g, "ms": 1e3 * min(times)}
        table.append(row)
        if progress:
            progress(row)
    best = min(table, key=lambda r: r["ms"])
    default = next(r for r in table if (r["resolve_grid"], r["min_group_log2"]) == DEFAULT)
    return {"workload": f"{L}x{L} kicked Ising, {steps} Trotter steps, theta_h {THETA_H} at once (own engines "
                        "and streams), cutoff 10, lower_atol 1e-5",
            "best": best, "default": default, "table": table}


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--lattice", type=int, default=10)
    ap.add_argument("--trotter-steps", type=int, default=20)
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("--out", default=str(_engine.TUNING_FILE))
    args = ap.parse_args()
    res = tune(args.lattice, args.trotter_steps, args.repeats,
               progress=lambda r: print(json.dumps(r), file=sys.stderr, flush=True))
    Path(args.out).write_text(json.dumps(res, indent=1) + "\n")
    print(json.dumps({"best": res["best"], "default": res["default"]}))


if __name__ == "__main__":
    main()
