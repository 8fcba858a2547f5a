"""monoprop_amd.concurrent.run_concurrently on the host: results in task order, the worker bound, and
failure reporting (every task still runs; the first failure in task order is raised)."""

from __future__ import annotations

import threading
import time

import pytest

from monoprop_amd.concurrent import run_concurrently


def test_results_in_task_order():
    def task(i):
        time.sleep(0.01 * (5 - i))  # later tasks finish first
        return i * i

    assert run_concurrently([lambda i=i: task(i) for i in range(5)]) == [0, 1, 4, 9, 16]
    assert run_concurrently([]) == []


def test_worker_bound():
    lock, live, peak = threading.Lock(), [0], [0]

    def task():
        with lock:
            live[0] += 1
            peak[0] = max(peak[0], live[0])
        time.sleep(0.02)
        with lock:
            live[0] -= 1

    run_concurrently([task] * 8, max_workers=3)
    assert 1 <= peak[0] <= 3


def test_first_failure_raised_after_all_ran():
    ran = []

    def ok(i):
        time.sleep(0.01)
        ran.append(i)
        return i

    def bad(msg):
        raise ValueError(msg)

    with pytest.raises(ValueError, match="first"):
        run_concurrently([lambda: ok(0), lambda: bad("first"), lambda: ok(2), lambda: bad("second")])
    assert sorted(ran) == [0, 2]
