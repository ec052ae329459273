"""GPU parity on the benchmark configurations' own frames.

1080p gray8 (lossless, clamped gradient), one 4K RGB 8-bit q90 frame (forced
order-0), and 4K gray10 / gray16 frames at the reachable steps of the
high-bit-depth configuration (gray10 q100 / q91, gray16 q100 / q87, where q87
may take the refined step): streams byte-identical to the oracle's and
decodes identical to the oracle's reconstruction.
"""

import os

import numpy as np
import pytest
import torch

from fastvid_amd import synthetic

pytestmark = pytest.mark.gpu


def _threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture(scope="module")
def fc():
    import fastvid_cuda
    return fastvid_cuda


def _check(fc, oracle, planes, layout, bit_depth, quality):
    dev = tuple(torch.from_numpy(p.astype(np.int32)).to(torch.uint16).cuda() for p in planes)
    stream = fc.encode(dev, layout=layout, bit_depth=bit_depth, quality=quality)
    got = bytes(stream.cpu().numpy().tobytes())
    ref = oracle.encode(planes, layout=layout, bit_depth=bit_depth, quality=quality,
                        threads=_threads())
    assert len(got) == len(ref)
    assert got == ref
    decoded = fc.decode(stream)
    expected = oracle.decode(ref, threads=_threads())
    for a, b in zip(decoded, expected):
        assert np.array_equal(a.cpu().numpy(), b)
    return got


def test_config0_gray8_1080p(fc, oracle):
    plane = synthetic.generate(0, 1920, 1080, 8, 3.0)
    got = _check(fc, oracle, (plane,), "gray", 8, 90)
    decoded = fc.decode(torch.frombuffer(bytearray(got), dtype=torch.uint8))
    assert np.array_equal(decoded[0].cpu().numpy(), plane)


def test_config1_rgb8_4k_frame(fc, oracle):
    _check(fc, oracle, synthetic.rgb_frame(1, 3840, 2160), "rgb444", 8, 90)


@pytest.mark.parametrize("bit_depth,quality", [(10, 100), (10, 91), (16, 100), (16, 87)])
def test_config2_high_bit_depth_4k(fc, oracle, bit_depth, quality):
    sigma = 12.0 if bit_depth == 10 else 200.0
    plane = synthetic.generate(1, 3840, 2160, bit_depth, sigma)
    _check(fc, oracle, (plane,), "gray", bit_depth, quality)
