"""Build the HIP engine library for gfx950 (in-tree: ``monoprop_amd/libmonoprop_amd.so``).

Flags:
  ``-ffp-contract=off``   keep the reference's unfused FP64 operation order;
  ``-Xclang -target-feature -Xclang -packed-fp32-ops``   no packed FP32 instructions;
  ``--offload-arch=gfx950``   MI355X only.
"""

from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libmonoprop_amd.so"
SOURCES = [CSRC / "engine.hip"]
HEADERS = [CSRC / h for h in ("algebra.h", "device_common.h", "kernels_pergate.h", "kernels_fused.h",
                              "kernels_reduce_graph.h", "engine_host.h")] + [HERE.parent / "include" / "monoprop_amd.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def build(force: bool = True) -> Path:
    """Compile the engine; with ``force=False`` reuse a library newer than every source."""
    if not force and LIB.exists():
        newest = max(p.stat().st_mtime for p in SOURCES + HEADERS)
        if LIB.stat().st_mtime >= newest:
            return LIB
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, *map(str, SOURCES), "-o", str(tmp)]
    subprocess.run(cmd, check=True)
    tmp.replace(LIB)
    return LIB


if __name__ == "__main__":
    print(build())
