"""Seeded synthetic frames for the benchmark workloads (numpy, host side).

Generator G(seed, W, H, bit_depth, sigma), in 8-bit units before scaling:
a linear ramp 0..255 along a random direction, plus three 2-D sinusoids
(period U(64, 512) px, random direction and phase, amplitude 20), plus
Gaussian noise (sigma given in units of the target bit depth), then 20
axis-aligned rectangles (position uniform over the frame, sides U{16..400})
overwritten with a uniform random intensity; finally scaled by
(2^bd - 1)/255, rounded and clipped.  Sinusoids are evaluated as outer
products of 1-D tables, so a 4K plane costs a few vector passes.
"""

from __future__ import annotations

import numpy as np

RESOLUTIONS = ((640, 360), (1280, 720), (1920, 1080), (2560, 1440), (3840, 2160))


def _smooth_field(rng: np.random.Generator, width: int, height: int) -> np.ndarray:
    x = np.arange(width, dtype=np.float64)
    y = np.arange(height, dtype=np.float64)
    angle = rng.uniform(0.0, 2.0 * np.pi)
    ramp_x = np.cos(angle) * x / width
    ramp_y = np.sin(angle) * y / height
    lo = min(ramp_x.min(), ramp_x.max()) + min(ramp_y.min(), ramp_y.max())
    hi = max(ramp_x.min(), ramp_x.max()) + max(ramp_y.min(), ramp_y.max())
    scale = 255.0 / (hi - lo) if hi > lo else 0.0
    img = ((ramp_y[:, None] - lo) * scale + ramp_x[None, :] * scale).astype(np.float32)
    for _ in range(3):
        period = rng.uniform(64.0, 512.0)
        direction = rng.uniform(0.0, 2.0 * np.pi)
        phase = rng.uniform(0.0, 2.0 * np.pi)
        k = 2.0 * np.pi / period
        ax = k * np.cos(direction) * x + phase
        ay = k * np.sin(direction) * y
        img += (20.0 * np.cos(ay)[:, None] * np.sin(ax)[None, :]).astype(np.float32)
        img += (20.0 * np.sin(ay)[:, None] * np.cos(ax)[None, :]).astype(np.float32)
    return img


def _rectangles(rng: np.random.Generator, img: np.ndarray) -> None:
    height, width = img.shape
    for _ in range(20):
        x0 = int(rng.integers(0, width))
        y0 = int(rng.integers(0, height))
        w = int(rng.integers(16, 401))
        h = int(rng.integers(16, 401))
        img[y0:y0 + h, x0:x0 + w] = float(rng.integers(0, 256))


def _finish(img: np.ndarray, bit_depth: int) -> np.ndarray:
    maximum = (1 << bit_depth) - 1
    out = np.rint(img * np.float32(maximum / 255.0))
    return np.clip(out, 0, maximum).astype(np.uint16)


def generate(seed: int, width: int, height: int, bit_depth: int = 8, sigma: float = 3.0) -> np.ndarray:
    """One plane of G(seed, W, H, bit_depth, sigma) as uint16."""
    rng = np.random.default_rng(seed)
    img = _smooth_field(rng, width, height)
    if sigma > 0:
        scale = 255.0 / ((1 << bit_depth) - 1)
        img += rng.standard_normal((height, width), dtype=np.float32) * np.float32(sigma * scale)
    _rectangles(rng, img)
    return _finish(img, bit_depth)


def rgb_frame(seed: int, width: int = 3840, height: int = 2160) -> tuple:
    """RGB 4:4:4 8-bit frame: shared base G(seed, W, H, 8, 0), per channel an
    N(0, 15) offset and independent N(0, 2) noise."""
    rng = np.random.default_rng(seed)
    base = _smooth_field(rng, width, height)
    _rectangles(rng, base)
    planes = []
    for _ in range(3):
        offset = np.float32(rng.normal(0.0, 15.0))
        noise = rng.standard_normal((height, width), dtype=np.float32) * np.float32(2.0)
        planes.append(np.clip(np.rint(base + offset + noise), 0, 255).astype(np.uint16))
    return tuple(planes)


# quality giving each quantiser step of the mixed batch (quantization_step
# of cuda/csrc/encode.cu:71-85): gray8 is always lossless.
STEP_QUALITY = {
    ("gray", 8): {1: 100, 3: 100, 7: 100},
    ("rgb444", 8): {1: 100, 3: 90, 7: 66},
    ("gray", 10): {1: 100, 3: 97, 7: 91},
}


def mixed_batch_specs(seed: int = 3, count: int = 64):
    """Shapes, formats and steps of the mixed batch: 64 frames drawn from five
    resolutions and {gray8, rgb444-8, gray10}, steps from {1, 3, 7}."""
    rng = np.random.default_rng(seed)
    formats = (("gray", 8), ("rgb444", 8), ("gray", 10))
    specs = []
    for i in range(count):
        w, h = RESOLUTIONS[int(rng.integers(0, len(RESOLUTIONS)))]
        layout, bd = formats[int(rng.integers(0, 3))]
        step = (1, 3, 7)[int(rng.integers(0, 3))]
        specs.append({"index": i, "width": w, "height": h, "layout": layout, "bit_depth": bd,
                      "step": step, "quality": STEP_QUALITY[(layout, bd)][step],
                      "seed": seed * 1000 + i})
    return specs


def spec_planes(spec: dict) -> tuple:
    w, h, bd = spec["width"], spec["height"], spec["bit_depth"]
    if spec["layout"] == "rgb444":
        return rgb_frame(spec["seed"], w, h)
    sigma = 3.0 if bd == 8 else 12.0
    return (generate(spec["seed"], w, h, bd, sigma),)


# ------------------------------------------------- device-generated frames

def _ramp_params(rng: np.random.Generator, width: int, height: int):
    angle = rng.uniform(0.0, 2.0 * np.pi)
    cx, cy = np.cos(angle), np.sin(angle)
    rx = (0.0, cx * (width - 1) / width)
    ry = (0.0, cy * (height - 1) / height)
    lo, hi = min(rx) + min(ry), max(rx) + max(ry)
    scale = 255.0 / (hi - lo) if hi > lo else 0.0
    return cx, cy, lo, scale


def _device_params(rng: np.random.Generator, width: int, height: int, bit_depth: int):
    """Ramp, sinusoid and rectangle parameters of one frame (host draws)."""
    import ctypes

    from . import _native
    p = _native.SynthParams()
    p.width, p.height, p.pitch, p.bit_depth = width, height, width, bit_depth
    p.ramp_cx, p.ramp_cy, p.ramp_lo, p.ramp_scale = _ramp_params(rng, width, height)
    for s in range(3):
        period = rng.uniform(64.0, 512.0)
        direction = rng.uniform(0.0, 2.0 * np.pi)
        k = 2.0 * np.pi / period
        p.kx[s], p.ky[s] = k * np.cos(direction), k * np.sin(direction)
        p.phase[s] = rng.uniform(0.0, 2.0 * np.pi)
    p.rect_count = 20
    rects = np.stack([rng.integers(0, width, 20), rng.integers(0, height, 20),
                      rng.integers(16, 401, 20), rng.integers(16, 401, 20),
                      rng.integers(0, 256, 20)], axis=1).astype(np.int32)
    ctypes.memmove(p.rects, rects.ctypes.data, rects.nbytes)
    return p


def device_frame_params(seed: int, width: int, height: int, layout: str,
                        bit_depth: int) -> list:
    """Generator parameters of each plane of ``device_frame(seed, ...)``."""
    from . import _native
    rng = np.random.default_rng(seed)
    base = _device_params(rng, width, height, bit_depth)
    planes = []
    if layout == "rgb444":
        channels = []
        for _ in range(3):
            channels.append((float(rng.normal(0.0, 15.0)), int(rng.integers(0, 1 << 63))))
        for offset, noise_seed in channels:
            p = _native.SynthParams.from_buffer_copy(base)
            p.offset, p.noise_sigma, p.noise_seed, p.rects_under_noise = offset, 2.0, noise_seed, 1
            planes.append(p)
    elif layout == "gray":
        sigma = 3.0 if bit_depth == 8 else 12.0
        base.noise_sigma = sigma * 255.0 / ((1 << bit_depth) - 1)
        base.noise_seed = int(rng.integers(0, 1 << 63))
        planes.append(base)
    else:
        raise ValueError("device frames are gray or rgb444")
    return planes


def launch_device_frame(params: list, device, stream=None) -> tuple:
    """Allocates the planes and enqueues their generation on `stream`
    (default: the current stream); the planes are ready once it reaches
    this point."""
    import ctypes

    import torch

    from . import _native
    if stream is None:
        stream = torch.cuda.current_stream(device)
    handle = ctypes.c_void_p(stream.cuda_stream)
    out = []
    with torch.cuda.stream(stream):
        for p in params:
            t = torch.empty((int(p.height), int(p.width)), dtype=torch.uint16, device=device)
            rc = _native.lib().fvh_synth_plane(ctypes.byref(p), ctypes.c_void_p(t.data_ptr()),
                                               handle)
            if rc != 0:
                raise RuntimeError(f"fvh_synth_plane failed ({rc})")
            out.append(t)
    return tuple(out)


def device_frame(seed: int, width: int, height: int, layout: str, bit_depth: int, device,
                 stream=None) -> tuple:
    """A frame of the serving workload generated on the device from `seed`.

    Same formula as ``generate`` (gray: field + noise, then rectangles; sigma
    3 at 8 bits, 12 at higher depths, in units of the depth) and
    ``rgb_frame`` (shared field + rectangles, per channel an N(0, 15) offset
    and N(0, 2) noise), but the noise comes from a counter-based generator on
    the device, so the frame is not the host generator's frame of the same
    seed.  Enqueued on `stream` (default: the current stream)."""
    return launch_device_frame(device_frame_params(seed, width, height, layout, bit_depth),
                               device, stream)


def frame(seed: int, width: int, height: int, layout: str, bit_depth: int) -> tuple:
    """Host planes of one frame in any layout at any depth (the evaluator
    workloads): gray is G(seed) with sigma 3 at 8 bits and 12 above; RGB 4:4:4
    is ``rgb_frame`` rescaled to the depth; YUV 4:2:2 is a luma plane
    G(seed) and two chroma planes G(seed + 1), G(seed + 2) of width
    ceil(W/2)."""
    sigma = 3.0 if bit_depth == 8 else 12.0
    if layout == "gray":
        return (generate(seed, width, height, bit_depth, sigma),)
    if layout == "yuv422":
        cw = (width + 1) // 2
        return (generate(seed, width, height, bit_depth, sigma),
                generate(seed + 1, cw, height, bit_depth, sigma),
                generate(seed + 2, cw, height, bit_depth, sigma))
    if layout == "rgb444":
        planes = rgb_frame(seed, width, height)
        if bit_depth == 8:
            return planes
        maximum = (1 << bit_depth) - 1
        return tuple(np.clip(np.rint(p.astype(np.float32) * np.float32(maximum / 255.0)), 0,
                             maximum).astype(np.uint16) for p in planes)
    raise ValueError(f"unsupported layout {layout!r}")
