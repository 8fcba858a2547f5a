"""Independent propagations at once on one GPU.

Every propagator owns one engine, and every engine its own HIP stream; the engine library is called
with the interpreter lock released (ctypes), so propagators driven from different host threads run
their launches side by side on the device.  One propagation is a chain of short, dependent launches
(per gate a sweep over the store, then a resolve that occupies only as many blocks as the gate has
256-emitter chunks), which leaves most of the GPU idle; the launches of other propagations fill it.
"""

from __future__ import annotations

from collections.abc import Callable, Iterable
from concurrent.futures import ThreadPoolExecutor
from typing import Any

from . import _engine


def run_concurrently(tasks: Iterable[Callable[[], Any]], max_workers: int | None = None) -> list:
    """Call each task (a callable without arguments, typically one that creates a propagator and
    propagates it) on a host thread of its own, at most ``max_workers`` at once (default: all); return
    their results in task order.  Tasks must not share a propagator.  If a task raises, the others still
    run to completion and the first failure (in task order) is raised."""
    tasks = list(tasks)
    if not tasks:
        return []
    _engine.load_library()
    _engine.default_launch()
    workers = len(tasks) if max_workers is None else max(1, min(int(max_workers), len(tasks)))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        futures = [pool.submit(t) for t in tasks]
        done = [f.exception() for f in futures]
    for f, err in zip(futures, done):
        if err is not None:
            raise err
    return [f.result() for f in futures]
