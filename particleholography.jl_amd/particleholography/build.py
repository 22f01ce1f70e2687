"""Build libholo.so in-tree (hipcc for gfx950) through the Makefile next to the sources."""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

CSRC = Path(__file__).resolve().parent.parent / "csrc"


def build(force: bool = True, jobs: int = 8) -> Path:
    """Compile libholo.so; returns its path. ``force`` rebuilds even when up to date."""
    env = dict(os.environ)
    cmd = ["make", "-C", str(CSRC), f"-j{jobs}"]
    if force:
        cmd.append("-B")
    subprocess.run(cmd, check=True, env=env)
    return CSRC.parent / "lib" / "libholo.so"


if __name__ == "__main__":  # pragma: no cover
    print(build())
