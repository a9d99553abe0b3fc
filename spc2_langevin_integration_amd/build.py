"""Compiles the gfx950 HIP library (and, when asked, the host C++ drop-in) in-tree."""
from __future__ import annotations

import os
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")


def build(jobs: int = 8, force: bool = True) -> str:
    """Runs the csrc Makefile (hipcc --offload-arch=gfx950).  Returns the .so path."""
    cmd = ["make", "-C", CSRC, f"-j{jobs}"]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)
    return os.path.join(PKG_DIR, "libspc2_hip.so")
