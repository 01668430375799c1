"""In-tree build of librgf_hip.so and RGFSolver (hipcc --offload-arch=gfx950)."""
from __future__ import annotations

import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 6) -> None:
    """Compile every HIP translation unit for gfx950 and link the library and CLI."""
    csrc = os.path.join(_HERE, "csrc")
    subprocess.run(["make", "-s", "-C", csrc, f"-j{jobs}", "-B"], check=True)
