"""Writes generator_digests.json: SHA-256 of the synthetic workload frames.

The workloads' synthetic frames (fastvid_amd.synthetic) are this package's own
definition; their digests freeze it, so a change to the generator (or to
numpy's PCG64 streams) shows up as a test failure instead of silently
changing every benchmark input.  Run from the repository root:

    python tests/golden/make_generator_digests.py
"""

from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from fastvid_amd import synthetic  # noqa: E402


def digest(planes) -> str:
    h = hashlib.sha256()
    for p in planes:
        h.update(str(p.shape).encode())
        h.update(p.astype("<u2").tobytes())
    return h.hexdigest()


def cases():
    """name -> planes, small enough to regenerate in a test."""
    yield "config0_gray8_1920x1080_seed0", (synthetic.generate(0, 1920, 1080, 8, 3.0),)
    yield "config1_rgb8_3840x2160_seed1", synthetic.rgb_frame(1, 3840, 2160)
    yield "config2_gray10_3840x2160_seed1", (synthetic.generate(1, 3840, 2160, 10, 12.0),)
    yield "config2_gray16_3840x2160_seed1", (synthetic.generate(1, 3840, 2160, 16, 200.0),)
    for spec in synthetic.mixed_batch_specs()[:4]:
        yield f"config3_frame{spec['index']}", synthetic.spec_planes(spec)


def main() -> int:
    out = {name: digest(planes) for name, planes in cases()}
    path = Path(__file__).with_name("generator_digests.json")
    path.write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")
    print(f"wrote {len(out)} digests to {path}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
