"""Shared test inputs.

``source_planes`` restates the reference test generator
(cuda/tests/test_codec.py:9-30) in numpy: an exact integer formula per plane.
"""

from __future__ import annotations

import struct

import numpy as np


def source_planes(layout, width, height, bit_depth, pattern="polynomial"):
    maximum = (1 << bit_depth) - 1
    if layout == "gray":
        widths = [width]
    else:
        cw = (width + 1) // 2 if layout == "yuv422" else width
        widths = [width, cw, cw]
    planes = []
    for seed, plane_width in enumerate(widths, 1):
        index = np.arange(plane_width * height, dtype=np.int64)
        if pattern == "zero":
            values = np.zeros_like(index)
        elif pattern == "ramp":
            values = index * (seed * 2 + 1) + seed * 17
        else:
            values = index * (977 + seed * 13) + index * index * 17 + seed * 131
        planes.append((values % (maximum + 1)).astype(np.uint16).reshape(height, plane_width))
    return tuple(planes)


def shard_modes(stream: bytes) -> set:
    """Mode byte of every shard record (cuda/tests/test_codec.py:33-49)."""
    tile_count = struct.unpack_from("<I", stream, 28)[0]
    modes = set()
    for tile in range(tile_count):
        entry = 32 + tile * 32
        width, height = struct.unpack_from("<II", stream, entry + 12)
        cursor = struct.unpack_from("<Q", stream, entry + 20)[0]
        tile_end = cursor + struct.unpack_from("<I", stream, entry + 28)[0]
        decoded = 0
        while decoded < width * height:
            mode = stream[cursor]
            body_length = struct.unpack_from("<H", stream, cursor + 1)[0]
            modes.add(mode)
            cursor += 3 + body_length
            decoded += min(4096, width * height - decoded)
        assert cursor == tile_end
    return modes


def noisy_planes(layout, width, height, bit_depth, seed, sigma=None):
    """Smooth field + noise + flat rectangles: exercises every entropy mode."""
    rng = np.random.default_rng(seed)
    maximum = (1 << bit_depth) - 1
    if layout == "gray":
        widths = [width]
    else:
        cw = (width + 1) // 2 if layout == "yuv422" else width
        widths = [width, cw, cw]
    sigma = maximum / 80.0 if sigma is None else sigma
    planes = []
    for pw in widths:
        yy, xx = np.mgrid[0:height, 0:pw]
        base = (np.sin(xx / (7 + rng.random() * 20)) + np.cos(yy / (5 + rng.random() * 30))) * maximum / 5
        base += maximum / 2 + rng.normal(0, sigma, size=base.shape)
        for _ in range(3):
            x0, y0 = rng.integers(0, pw), rng.integers(0, height)
            base[y0:y0 + rng.integers(1, 40), x0:x0 + rng.integers(1, 40)] = rng.integers(0, maximum + 1)
        planes.append(np.clip(np.rint(base), 0, maximum).astype(np.uint16))
    return tuple(planes)


def lattice_plane(width, height, step, seed=0, density=0.02):
    """16-bit plane whose samples all lie on multiples of `step`: a smooth
    8-bit field (generator G, no noise) divided by 8, plus sparse unit
    spikes, times `step`.  Encoded at the quality whose refined gray16 step
    equals `step`, the refined pass codes the lattice exactly and its payload
    drops below the baseline pass's (the gray16 probe is kept)."""
    from fastvid_amd import synthetic
    rng = np.random.default_rng(seed)
    spikes = (rng.random((height, width)) < density).astype(np.int64)
    base = synthetic.generate(5, width, height, 8, 0.0).astype(np.int64) // 8
    return np.clip((base + spikes) * step, 0, 65535).astype(np.uint16)
