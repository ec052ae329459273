"""Batched, graph-captured forms of the codec.

``GraphEncoder`` plans a fixed list of frame shapes once, keeps its input
planes, workspace and output resident on the device, and captures the whole
encode launch sequence (planning kernels, predictor, analysis, rANS,
assembly, emission: one launch sequence for the batch) in a graph, so a batch
costs one graph replay plus one result read-back.  ``GraphDecoder`` does the
same for decode with the stream sizes in device memory and every header read
on the device (``fvh_decode_launch_sized``), so new streams of the planned
geometry need no host parsing.  ``autotune_tile_size`` picks the tile
geometry that encodes a given frame format fastest.

The reference has no batched form: its batch is a Python loop over frames
(fastvid/__init__.py:44-53).
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass
from typing import Sequence

import numpy as np
import torch

from . import _native
from .codec import EncodeJob, _read_header, _read_results, frame_args


@dataclass(frozen=True)
class FrameSpec:
    """Shape and coding parameters of one frame of a batch."""
    layout: str
    width: int
    height: int
    bit_depth: int
    quality: int
    tile_size: tuple = (256, 128)
    frame_rate: tuple = (24, 1)

    def plane_shapes(self) -> list:
        if self.layout == "gray":
            return [(self.height, self.width)]
        cw = (self.width + 1) // 2 if self.layout == "yuv422" else self.width
        return [(self.height, self.width), (self.height, cw), (self.height, cw)]

    def luma_pixels(self) -> int:
        return self.width * self.height


def _capture(launch, stream: torch.cuda.Stream) -> torch.cuda.CUDAGraph:
    """Runs ``launch`` once eagerly (first-use initialisation happens outside
    the capture), then captures it."""
    with torch.cuda.stream(stream):
        launch()
    stream.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=stream):
        launch()
    stream.synchronize()
    return graph


class GraphEncoder:
    """Encodes batches of frames of fixed shapes with one graph replay.

    ``inputs[i]`` are the resident uint16 planes of frame i: write frames
    into them (or pass them to ``encode``), then ``run()`` returns the
    streams as views into the resident output, valid until the next run."""

    def __init__(self, specs: Sequence[FrameSpec], device=None, graph: bool = True):
        if not specs:
            raise ValueError("no frames to encode")
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        self.specs = list(specs)
        self.inputs = [tuple(torch.zeros(shape, dtype=torch.uint16, device=self.device)
                             for shape in s.plane_shapes()) for s in self.specs]
        args = [frame_args(planes, layout=s.layout, bit_depth=s.bit_depth, quality=s.quality,
                           frame_rate=s.frame_rate, tile_size=s.tile_size)
                for planes, s in zip(self.inputs, self.specs)]
        with torch.cuda.device(self.device):
            self.job = EncodeJob(args).allocate()
            self.stream = torch.cuda.Stream(self.device)
            self.graph = _capture(lambda: self.job.launch(stream=self.stream),
                                  self.stream) if graph else None

    def resident_bytes(self) -> int:
        planes = sum(p.numel() * 2 for f in self.inputs for p in f)
        return planes + self.job.workspace.numel() + self.job.output.numel()

    def load(self, frames: Sequence) -> None:
        if len(frames) != len(self.inputs):
            raise ValueError("frame count does not match the planned batch")
        for dst, src in zip(self.inputs, frames):
            src = (src,) if isinstance(src, torch.Tensor) else tuple(src)
            if len(src) != len(dst):
                raise ValueError("plane count does not match the planned frame")
            for d, p in zip(dst, src):
                if tuple(p.shape) != tuple(d.shape):
                    raise ValueError("plane shape does not match the planned frame")
                d.copy_(p, non_blocking=True)

    def launch(self) -> None:
        """Enqueues the batch on the encoder's stream (after the current
        stream's pending work, e.g. ``load``)."""
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        if self.graph is not None:
            with torch.cuda.stream(self.stream):
                self.graph.replay()
        else:
            self.job.launch(stream=self.stream)

    def finish(self) -> list:
        self.stream.synchronize()
        return self.job.finish()

    def run(self) -> list:
        self.launch()
        return self.finish()

    def encode(self, frames: Sequence) -> list:
        self.load(frames)
        return self.run()


class GraphDecoder:
    """Decodes batches of streams of fixed geometry with one graph replay.

    ``templates`` are streams (or their headers) giving each frame's
    geometry; ``capacities[i]`` bounds the size of stream i (default: twice
    the template's size).  ``decode(streams)`` copies the streams and their
    sizes into resident buffers and replays; outputs are views into resident
    planes, valid until the next run."""

    def __init__(self, templates: Sequence[torch.Tensor], capacities=None, device=None,
                 graph: bool = True, wavefront: bool = True):
        if not templates:
            raise ValueError("no streams to decode")
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        n = len(templates)
        self.infos = (_native.StreamInfo * n)()
        for i, t in enumerate(templates):
            self.infos[i] = _read_header(t)
        caps = capacities if capacities is not None else [
            2 * int(self.infos[i].stream_size) + 4096 for i in range(n)]
        for i, c in enumerate(caps):
            if int(c) < 32 + 32 * int(self.infos[i].tile_count):
                raise ValueError("stream capacity is smaller than the header and directory")
        lib = _native.lib()
        msg = _native.message_buffer()
        ws = ctypes.c_uint64()
        _native.check(lib.fvh_decode_plan(ctypes.addressof(self.infos), n, ctypes.byref(ws), msg,
                                          len(msg)), msg)
        with torch.cuda.device(self.device):
            self.streams = [torch.zeros(int(c), dtype=torch.uint8, device=self.device)
                            for c in caps]
            self.sizes = torch.zeros(n, dtype=torch.int64, device=self.device)
            self.workspace = torch.empty(max(ws.value, 1), dtype=torch.uint8, device=self.device)
            self.outputs = [torch.empty(int(self.infos[i].total_samples), dtype=torch.uint16,
                                        device=self.device) for i in range(n)]
            self.results = torch.empty(n * _native.RESULT_BYTES, dtype=torch.uint8,
                                       device=self.device)
            self._ins = (ctypes.c_void_p * n)(*[t.data_ptr() for t in self.streams])
            self._outs = (ctypes.c_void_p * n)(*[t.data_ptr() for t in self.outputs])
            self._frames = [self.planes(i) for i in range(n)]
            self.wavefront = bool(wavefront)
            self.stream = torch.cuda.Stream(self.device)
            self.graph = _capture(self._launch_now, self.stream) if graph else None

    def _launch_now(self) -> None:
        msg = _native.message_buffer()
        _native.check(_native.lib().fvh_decode_launch_sized(
            ctypes.addressof(self.infos), self._ins, ctypes.c_void_p(self.sizes.data_ptr()),
            self._outs, len(self.streams), self.workspace.data_ptr(), self.workspace.numel(),
            self.results.data_ptr(), int(self.wavefront),
            ctypes.c_void_p(self.stream.cuda_stream), msg, len(msg)), msg)

    def resident_bytes(self) -> int:
        return (sum(t.numel() for t in self.streams) + self.workspace.numel() +
                sum(t.numel() * 2 for t in self.outputs))

    def load(self, streams: Sequence[torch.Tensor]) -> None:
        if len(streams) != len(self.streams):
            raise ValueError("stream count does not match the planned batch")
        sizes = []
        for dst, src in zip(self.streams, streams):
            if src.dtype != torch.uint8 or src.dim() != 1:
                raise RuntimeError("encoded must be a one-dimensional uint8 tensor")
            if src.numel() > dst.numel():
                raise ValueError("stream is larger than its planned capacity")
            dst.narrow(0, 0, src.numel()).copy_(src, non_blocking=True)
            sizes.append(src.numel())
        self.sizes.copy_(torch.tensor(sizes, dtype=torch.int64), non_blocking=False)

    def launch(self) -> None:
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        if self.graph is not None:
            with torch.cuda.stream(self.stream):
                self.graph.replay()
        else:
            self._launch_now()

    def planes(self, i: int) -> tuple:
        info = self.infos[i]
        out = self.outputs[i]
        height, width = info.height, info.width
        planes = [out.narrow(0, 0, info.y_samples).view(height, width)]
        if info.layout != 0:
            sw = info.secondary_samples // height
            planes.append(out.narrow(0, info.y_samples, info.secondary_samples).view(height, sw))
            planes.append(out.narrow(0, info.y_samples + info.secondary_samples,
                                     info.secondary_samples).view(height, sw))
        return tuple(planes)

    def finish(self) -> list:
        status = _read_results(self.results, self.stream)["status"]
        bad = np.flatnonzero(status)
        if bad.size:
            first = int(status[bad[0]])
            if first == 12:
                raise RuntimeError("stream header does not match the planned geometry")
            raise RuntimeError(f"malformed entropy payload (CUDA status {first})")
        return list(self._frames)

    def decode(self, streams: Sequence[torch.Tensor]) -> list:
        self.load(streams)
        self.launch()
        return self.finish()


# ------------------------------------------------------------------ autotune

TILE_CANDIDATES = ((64, 64), (128, 64), (128, 128), (256, 64), (256, 128), (256, 256),
                   (512, 128))

_TUNED: dict = {}


def autotune_tile_size(layout: str, width: int, height: int, bit_depth: int, quality: int,
                       candidates=TILE_CANDIDATES, frames=None, reps: int = 3,
                       device=None) -> tuple:
    """Tile size (from ``candidates``) that encodes this frame format
    fastest, measured with graph replays on ``frames`` (default: one
    synthetic frame of the format).  Results are cached per format.  The
    stream stays a valid version-1 stream at the chosen tile size."""
    key = (layout, width, height, bit_depth, quality, tuple(candidates))
    if key in _TUNED:
        return _TUNED[key]
    if frames is None:
        from . import synthetic
        spec = {"layout": layout, "width": width, "height": height, "bit_depth": bit_depth,
                "seed": 12345}
        host = synthetic.spec_planes(spec)
        dev = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        frames = [tuple(torch.from_numpy(p.astype("int32")).to(torch.uint16).to(dev)
                        for p in host)]
    best, best_t = None, float("inf")
    for tile in candidates:
        spec = FrameSpec(layout, width, height, bit_depth, quality, tuple(tile))
        enc = GraphEncoder([spec] * len(frames), device=device)
        enc.load(frames)
        enc.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            enc.launch()
        enc.stream.synchronize()
        t = (time.perf_counter() - t0) / reps
        enc.finish()
        if t < best_t:
            best, best_t = tuple(tile), t
        del enc
    _TUNED[key] = best
    return best

