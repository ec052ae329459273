"""Workload inputs and the CPU leg of the benchmark configurations.

The synthetic frames are frozen by committed SHA-256 digests
(tests/golden/generator_digests.json, written by
tests/golden/make_generator_digests.py).  The 1080p gray8 configuration
(lossless, clamped-gradient predictor, per-shard mode choice) round-trips
bit-exactly through the oracle, as the reference's CPU-path check requires.
"""

import json
import os
import struct
from pathlib import Path

import numpy as np
import pytest

from fastvid_amd import synthetic

GOLDEN = Path(__file__).with_name("golden")


def _threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def test_generator_digests_are_frozen():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_digests",
                                                  GOLDEN / "make_generator_digests.py")
    module = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(module)
    expected = json.loads((GOLDEN / "generator_digests.json").read_text())
    got = {name: module.digest(planes) for name, planes in module.cases()}
    assert got == expected


def test_generator_ranges_and_shapes():
    g = synthetic.generate(0, 1920, 1080, 8, 3.0)
    assert g.shape == (1080, 1920) and g.dtype == np.uint16 and g.max() <= 255
    g16 = synthetic.generate(1, 640, 360, 16, 200.0)
    assert g16.max() <= 65535 and g16.max() > 50000
    r = synthetic.rgb_frame(1, 640, 360)
    assert len(r) == 3 and all(p.shape == (360, 640) and p.max() <= 255 for p in r)


def test_config0_gray8_round_trip_on_the_oracle(oracle):
    plane = synthetic.generate(0, 1920, 1080, 8, 3.0)
    stream = oracle.encode((plane,), layout="gray", bit_depth=8, quality=90,
                           threads=_threads())
    assert stream[:4] == b"FVID" and stream[4] == 1
    # gray8: step 1 at any quality (encode.cu:72-74), no MED flag
    assert stream[7] == 0
    assert oracle.quantization_step(0, 8, 90) == 1
    decoded = oracle.decode(stream, threads=_threads())
    assert np.array_equal(decoded[0], plane)
    # 8 x 9 tiles of 256x128
    assert struct.unpack_from("<I", stream, 28)[0] == 72


@pytest.mark.parametrize("bit_depth,quality,step", [(10, 100, 1), (10, 91, 7), (10, 92, 7),
                                                    (16, 100, 1)])
def test_config2_steps(oracle, bit_depth, quality, step):
    assert oracle.quantization_step(0, bit_depth, quality) == step


def test_config2_gray16_refined_step_is_417_at_q87(oracle):
    assert oracle.refined_gray16_step(87) == 417
    assert oracle.quantization_step(0, 16, 87) == 556
