"""Build recipe for the oracle library (test infrastructure only).

Compiles ``mp_oracle.cpp`` with g++ into ``oracle/liboracle_mp.so``.  ``-ffp-contract=off`` keeps the
floating-point operation order of the reference (no FMA contraction); ``-mpopcnt`` is the only ISA
extension used so the library also runs on the GPU host's CPUs; ``-fopenmp`` serves the optional
thread-parallel restatement (``orc_set_threads``), whose results equal the serial ones.

The reference itself cannot be built here (it needs C++23 ``<format>``/``<print>``, Boost >= 1.85,
hwloc and nanobind), so there is no ``oracle/_ref`` build.
"""

from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "mp_oracle.cpp"
LIB = HERE / "liboracle_mp.so"

CXXFLAGS = [
    "-O3",
    "-std=c++17",
    "-mpopcnt",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fopenmp",
    "-fPIC",
    "-shared",
    "-Wall",
]


def build(force: bool = True) -> Path:
    """Compile the oracle; with ``force=False`` reuse an up-to-date library."""
    if not force and LIB.exists() and LIB.stat().st_mtime >= SRC.stat().st_mtime:
        return LIB
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["g++", *CXXFLAGS, str(SRC), "-o", str(tmp)]
    subprocess.run(cmd, check=True)
    tmp.replace(LIB)
    return LIB


if __name__ == "__main__":
    print(build())
