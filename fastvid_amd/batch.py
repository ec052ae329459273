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
from .codec import EncodeJob, _pinned_results, _read_header, _read_results, frame_args


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
            self.graph = _capture(lambda: self.job.launch(stream=self.stream, readback=False),
                                  self.stream) if graph else None

    def resident_bytes(self) -> int:
        planes = sum(p.numel() * 2 for f in self.inputs for p in f)
        return planes + self.job.workspace.numel() + self.job.output.numel()

    def load(self, frames: Sequence) -> None:
        """Copies frames into the resident input planes on the current
        stream, after any replay still reading them (a ``launch()`` not yet
        finished)."""
        if len(frames) != len(self.inputs):
            raise ValueError("frame count does not match the planned batch")
        torch.cuda.current_stream(self.device).wait_stream(self.stream)
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
            self.job.launch(stream=self.stream, readback=False)

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
            self._host_results = _pinned_results(n)
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
        """Copies streams and their sizes into the resident buffers on the
        current stream, after any replay still reading them."""
        if len(streams) != len(self.streams):
            raise ValueError("stream count does not match the planned batch")
        torch.cuda.current_stream(self.device).wait_stream(self.stream)
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
        status = _read_results(self.results, self.stream, self._host_results)["status"]
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

# Tile geometries the tuner tries (the stream header carries the tile size, so
# any of them is a valid version-1 stream); 256x128 is the reference default.
DEFAULT_TILE = (256, 128)
TILE_CANDIDATES = ((32, 32), (64, 32), (64, 64), (128, 64), (128, 128), (256, 64), (256, 128),
                   (128, 256), (256, 256))
# Persistent band-wavefront waves per compute unit (16 lanes per tile, four
# tiles per wave; the library's default is 16).
WAVE_CANDIDATES = (4, 8, 12, 16, 24)


@dataclass
class TuneResult:
    """Outcome of ``autotune``: a tile size per frame format, the band waves
    per compute unit, and the measured batch times (seconds per replay of
    the objective) of the chosen and the default configuration."""
    tiles: dict
    band_waves: int
    seconds: float
    default_seconds: float
    evaluations: int
    log: list

    def tile_for(self, spec: "FrameSpec") -> tuple:
        return self.tiles.get(_group_key(spec), spec.tile_size)

    def apply(self, specs: Sequence["FrameSpec"]) -> list:
        return [FrameSpec(s.layout, s.width, s.height, s.bit_depth, s.quality,
                          self.tile_for(s), s.frame_rate) for s in specs]


def _group_key(spec: "FrameSpec") -> tuple:
    return (spec.layout, spec.width, spec.height, spec.bit_depth, spec.quality)


def _median_replay(enc: GraphEncoder, dec, reps: int) -> float:
    """Median of `reps` timings of one replay of the objective (encode, then
    decode when `dec` is given), HIP events on the replay streams."""
    times = []
    for _ in range(reps):
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record(enc.stream)
        enc.launch()
        if dec is not None:
            dec.stream.wait_stream(enc.stream)
            dec.launch()
            end.record(dec.stream)
        else:
            end.record(enc.stream)
        end.synchronize()
        times.append(start.elapsed_time(end) / 1e3)
    return float(np.median(times))


def autotune(specs: Sequence[FrameSpec], frames: Sequence, tile_candidates=TILE_CANDIDATES,
             wave_candidates=WAVE_CANDIDATES, reps: int = 5, objective: str = "encode+decode",
             per_format: bool = True, margin: float = 0.01, device=None,
             progress=None) -> TuneResult:
    """Tile geometry per frame format and the predictor's band waves per
    compute unit (the reconstruction keeps its own count) that minimise the
    graph replay time of the whole batch.

    The objective is one replay of a ``GraphEncoder`` over every frame of the
    batch (and, for "encode+decode", a ``GraphDecoder`` of its streams).  The
    search starts from the default (256x128 tiles, the library's wave count)
    and is greedy: the wave count over ``wave_candidates``, then one tile
    size for the whole batch, then (``per_format``) each format's tile size
    with the others held.  A change is kept only if it is at least
    ``margin`` faster; at the end the chosen and the default configuration
    are timed again alternately and the default is returned unless the
    chosen one is faster, so the result is never slower than the default
    on the batch it was tuned on.  The library's wave count is left at the
    chosen value.  ``progress``, if given, is called with each measurement's
    log entry."""
    from . import _native
    specs = list(specs)
    if len(frames) != len(specs):
        raise ValueError("one frame per spec")
    default_waves = _native.band_waves()
    groups = sorted({_group_key(s) for s in specs})
    log = []

    def measure(tiles: dict, waves: int) -> float:
        _native.band_waves(waves)
        tuned = [FrameSpec(s.layout, s.width, s.height, s.bit_depth, s.quality,
                           tiles.get(_group_key(s), DEFAULT_TILE), s.frame_rate) for s in specs]
        enc = GraphEncoder(tuned, device=device)
        enc.load(frames)
        streams = enc.run()
        dec = None
        if objective == "encode+decode":
            dec = GraphDecoder(streams, capacities=[int(t.numel()) for t in streams],
                               device=device)
            dec.load(streams)
            dec.launch()
            dec.finish()
        t = _median_replay(enc, dec, reps)
        log.append({"tiles": {"%s-%dx%d-%db-q%d" % k: list(v) for k, v in tiles.items()},
                    "band_waves": waves, "seconds": t})
        if progress is not None:
            progress(log[-1])
        del enc, dec
        return t

    base_tiles = {g: DEFAULT_TILE for g in groups}
    best_tiles, best_waves = dict(base_tiles), default_waves
    best = measure(best_tiles, best_waves)
    for w in wave_candidates:
        if w == best_waves:
            continue
        t = measure(best_tiles, w)
        if t < best * (1 - margin):
            best, best_waves = t, w
    for tile in tile_candidates:
        if tile == DEFAULT_TILE:
            continue
        trial = {g: tuple(tile) for g in groups}
        t = measure(trial, best_waves)
        if t < best * (1 - margin):
            best, best_tiles = t, trial
    if per_format and len(groups) > 1:
        for g in groups:
            for tile in tile_candidates:
                if tuple(tile) == best_tiles[g]:
                    continue
                trial = dict(best_tiles)
                trial[g] = tuple(tile)
                t = measure(trial, best_waves)
                if t < best * (1 - margin):
                    best, best_tiles = t, trial
    # the chosen configuration against the default, alternately
    chosen, default = [], []
    for _ in range(3):
        chosen.append(measure(best_tiles, best_waves))
        default.append(measure(base_tiles, default_waves))
    t_chosen, t_default = float(np.median(chosen)), float(np.median(default))
    if t_chosen >= t_default:
        best_tiles, best_waves, t_chosen = base_tiles, default_waves, t_default
    _native.band_waves(best_waves)
    return TuneResult(best_tiles, best_waves, t_chosen, t_default, len(log), log)


_TUNED: dict = {}


def autotune_tile_size(layout: str, width: int, height: int, bit_depth: int, quality: int,
                       candidates=TILE_CANDIDATES, frames=None, reps: int = 3,
                       device=None) -> tuple:
    """Tile size for one frame format alone (``autotune`` over a one-format
    batch of ``frames``, default one synthetic frame, tile size only).
    Results are cached per format."""
    key = (layout, width, height, bit_depth, quality, tuple(candidates))
    if key in _TUNED:
        return _TUNED[key]
    dev = torch.device(device) if device is not None else torch.device(
        "cuda", torch.cuda.current_device())
    if frames is None:
        from . import synthetic
        spec = {"layout": layout, "width": width, "height": height, "bit_depth": bit_depth,
                "seed": 12345}
        host = synthetic.spec_planes(spec)
        frames = [tuple(torch.from_numpy(p.astype("int32")).to(torch.uint16).to(dev)
                        for p in host)]
    from . import _native
    waves = _native.band_waves()
    result = autotune([FrameSpec(layout, width, height, bit_depth, quality)] * len(frames),
                      frames, tile_candidates=candidates, wave_candidates=(waves,), reps=reps,
                      objective="encode", per_format=False, device=dev)
    _TUNED[key] = result.tiles[(layout, width, height, bit_depth, quality)]
    return _TUNED[key]
