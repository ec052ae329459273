"""In-tree build of libfastvid_hip.so (hipcc, gfx950) through csrc/Makefile."""

from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent


def build(jobs: int = 8, force: bool = False) -> Path:
    """Compile every HIP source for gfx950; returns the library path."""
    env = dict(os.environ)
    env.setdefault("ARCH", "gfx950")
    args = ["make", "-s", "-C", str(HERE / "csrc"), f"-j{jobs}"]
    if force:
        subprocess.run(["make", "-s", "-C", str(HERE / "csrc"), "clean"], check=True, env=env)
    subprocess.run(args, check=True, env=env)
    return HERE / "_lib" / "libfastvid_hip.so"


if __name__ == "__main__":
    print(build())
