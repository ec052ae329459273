"""Evaluator-equivalent harness: the reference's correctness checks and
performance gates on synthetic frames, on MI355X.

    python -m fastvid_amd.evaluate [--output results.json] [--repetitions 20]

Runs the gate configurations of the reference evaluator
(scripts/evaluate.py:740-755) through a codec module with the public
``encode`` / ``decode`` / ``inspect`` API (default ``fastvid``):

  4K x 24 YUV 4:2:2 10-bit   throughput   encode >= 2.0, decode >= 3.0 GP/s
  4K x 24 RGB 4:4:4 10-bit   throughput   encode >= 1.5, decode >= 2.0 GP/s
  1080p RGB 4:4:4 10-bit      latency      encode < 1.0 ms, decode < 0.5 ms

at quality 90 and 256x128 tiles.  Per sample, as scripts/evaluate.py:535-629
does: two encodes of the first frame must be identical; ``inspect`` must
report the sample's width, height, format, bit depth, one frame and valid
metadata/overhead byte counts; decodes must return CUDA uint16 planes of
the sample's shapes, twice identical; a batch encode returns one stream per
frame and a batch decode one frame per stream.  Timing follows
scripts/evaluate.py:447-476: warm-ups, then repetitions each bracketed by
events on the current stream after a device synchronisation; the gate
statistic is the median (p95 and p99 are reported beside it).  GP/s counts
luma pixels.  The corpus, its SHA-256 manifests, the perceptual metrics
(libvship) and the edit-cycle robustness suite are not part of this harness:
the frames are the synthetic generator's (``synthetic.frame``) and no
quality gate is evaluated.
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import importlib
import json
import statistics
import sys
import time
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class Sample:
    id: str
    format: str
    bit_depth: int
    width: int
    height: int
    batch_frames: int

    @property
    def plane_shapes(self) -> tuple:
        cw = (self.width + 1) // 2 if self.format == "yuv422" else self.width
        if self.format == "gray":
            return ((self.height, self.width),)
        return ((self.height, self.width), (self.height, cw), (self.height, cw))

    @property
    def raw_bytes_per_frame(self) -> int:
        return sum(h * w for h, w in self.plane_shapes) * 2


GATE_SAMPLES = (
    Sample("4k-yuv422-10-x24", "yuv422", 10, 3840, 2160, 24),
    Sample("4k-rgb444-10-x24", "rgb444", 10, 3840, 2160, 24),
    Sample("1080p-rgb444-10", "rgb444", 10, 1920, 1080, 1),
)


def cuda_samples(operation, warmups: int, repetitions: int, torch):
    """(seconds per repetition, last result): events on the current stream
    around each repetition, after a device synchronisation."""
    for _ in range(warmups):
        operation()
    torch.cuda.synchronize()
    samples, result = [], None
    for _ in range(repetitions):
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record()
        result = operation()
        end.record()
        end.synchronize()
        samples.append(start.elapsed_time(end) / 1000.0)
    return samples, result


def distribution(seconds) -> dict:
    ordered = sorted(seconds)
    at = lambda p: ordered[min(len(ordered) - 1, int((len(ordered) - 1) * p))]  # noqa: E731
    return {"samples_seconds": list(seconds), "median_seconds": statistics.median(seconds),
            "min_seconds": ordered[0], "p95_seconds": at(0.95), "p99_seconds": at(0.99),
            "max_seconds": ordered[-1], "gate_statistic": "median"}


def performance_failures(result: dict) -> list:
    """The reference's gates (scripts/evaluate.py:740-755)."""
    failures = []
    perf, fmt = result["performance"], result["format"]
    size = (result["width"], result["height"])
    if perf["mode"] == "throughput" and result["frame_count"] == 24 and size == (3840, 2160):
        enc_min, dec_min = (2.0, 3.0) if fmt == "yuv422" else (1.5, 2.0)
        if perf["encode"]["gpps"] < enc_min:
            failures.append(f"encode throughput < {enc_min} GP/s")
        if perf["decode"]["gpps"] < dec_min:
            failures.append(f"decode throughput < {dec_min} GP/s")
    if perf["mode"] == "latency" and size == (1920, 1080) and fmt == "rgb444":
        if perf["encode"]["median_ms"] >= 1.0:
            failures.append("encode latency >= 1.0 ms")
        if perf["decode"]["median_ms"] >= 0.5:
            failures.append("decode latency >= 0.5 ms")
    return failures


def load_frames(sample: Sample, torch, seed: int = 1000):
    from . import synthetic
    seeds = [seed + 3 * i for i in range(sample.batch_frames)]
    with cf.ThreadPoolExecutor(max_workers=8) as pool:
        host = list(pool.map(lambda s: synthetic.frame(s, sample.width, sample.height,
                                                       sample.format, sample.bit_depth), seeds))
    return [tuple(torch.from_numpy(p.astype(np.int32)).to(torch.uint16).cuda() for p in f)
            for f in host]


def evaluate_sample(sample: Sample, codec, torch, args) -> dict:
    frames = load_frames(sample, torch)
    kw = dict(format=sample.format, bit_depth=sample.bit_depth, quality=args.quality,
              tile_size=tuple(args.tile_size))
    encode_one = lambda f: codec.encode(f, **kw)  # noqa: E731
    encode_batch = lambda b: codec.encode(b, **kw)  # noqa: E731

    # correctness, outside every timed region (scripts/evaluate.py:535-587)
    first = encode_one(frames[0])
    torch.cuda.synchronize()
    second = encode_one(frames[0])
    torch.cuda.synchronize()
    if not torch.equal(first, second):
        raise RuntimeError("two identical encodes produced different bitstreams")
    meta = codec.inspect(first)
    expected = {"width": sample.width, "height": sample.height, "format": sample.format,
                "bit_depth": sample.bit_depth, "frame_count": 1}
    wrong = {k: (meta.get(k), v) for k, v in expected.items() if meta.get(k) != v}
    if wrong:
        raise RuntimeError(f"bitstream metadata differs: {wrong}")
    for key in ("metadata_bytes", "container_overhead_bytes"):
        if key not in meta:
            raise RuntimeError(f"inspect() omitted required {key}")
    if not 0 <= int(meta["metadata_bytes"]) + int(meta["container_overhead_bytes"]) <= first.numel():
        raise RuntimeError("invalid metadata/container byte counts from inspect()")
    decoded = codec.decode(first)
    torch.cuda.synchronize()
    if tuple(tuple(p.shape) for p in decoded) != sample.plane_shapes:
        raise RuntimeError("decoded plane dimensions or frame format differ")
    if any(p.dtype != torch.uint16 or not p.is_cuda for p in decoded):
        raise RuntimeError("decoder did not return CUDA uint16 planes")
    again = codec.decode(first)
    torch.cuda.synchronize()
    if any(not torch.equal(a, b) for a, b in zip(decoded, again)):
        raise RuntimeError("repeated decode produced different pixels")
    if len(frames) == 1:
        streams = [first]
    else:
        streams = encode_batch(frames)
        if not isinstance(streams, (list, tuple)) or len(streams) != len(frames):
            raise RuntimeError("batch encode must return one independent stream per frame")
        if not torch.equal(streams[0], first):
            raise RuntimeError("batch encode differs from the single-frame encode")
    torch.cuda.synchronize()
    sizes = [int(s.numel()) for s in streams]
    pixels = sample.width * sample.height * len(frames)

    # timing (scripts/evaluate.py:597-629)
    enc_op = (lambda: encode_one(frames[0])) if len(frames) == 1 else (lambda: encode_batch(frames))
    enc_t, _ = cuda_samples(enc_op, args.warmups, args.repetitions, torch)
    dec_op = (lambda: codec.decode(streams[0])) if len(streams) == 1 else (
        lambda: codec.decode(streams))
    dec_t, _ = cuda_samples(dec_op, args.warmups, args.repetitions, torch)
    enc, dec = distribution(enc_t), distribution(dec_t)
    mode = "throughput" if len(frames) > 1 else "latency"
    if mode == "throughput":
        enc["gpps"] = pixels / enc["median_seconds"] / 1e9
        dec["gpps"] = pixels / dec["median_seconds"] / 1e9
    else:
        enc["median_ms"] = enc["median_seconds"] * 1e3
        dec["median_ms"] = dec["median_seconds"] * 1e3

    out = [decoded] if len(frames) == 1 else codec.decode(streams)
    torch.cuda.synchronize()
    if not isinstance(out, (list, tuple)) or len(out) != len(frames):
        raise RuntimeError("batch decode must return one frame per independent stream")
    for n, f in enumerate(out):
        if tuple(tuple(p.shape) for p in f) != sample.plane_shapes:
            raise RuntimeError(f"decoded frame {n} dimensions or format differ")
        if any(p.dtype != torch.uint16 or not p.is_cuda for p in f):
            raise RuntimeError(f"decoded frame {n} is not CUDA uint16")
    raw = sample.raw_bytes_per_frame * len(frames)
    return {
        "id": sample.id, "format": sample.format, "bit_depth": sample.bit_depth,
        "width": sample.width, "height": sample.height, "frame_count": len(frames),
        "correctness": {"passed": True, "deterministic": True},
        "compression": {"raw_bytes": raw, "encoded_bytes": sum(sizes),
                        "ratio": raw / sum(sizes),
                        "bits_per_luma_pixel": sum(sizes) * 8 / pixels,
                        "metadata_bytes_first_frame": int(meta["metadata_bytes"])},
        "performance": {"mode": mode, "warmups": args.warmups, "repetitions": args.repetitions,
                        "encode": enc, "decode": dec},
    }


def summary_line(result: dict) -> dict:
    perf = result["performance"]
    line = {"id": result["id"], "mode": perf["mode"]}
    for side in ("encode", "decode"):
        d = perf[side]
        line[side] = {"median_ms": round(d["median_seconds"] * 1e3, 3),
                      "p95_ms": round(d["p95_seconds"] * 1e3, 3),
                      "p99_ms": round(d["p99_seconds"] * 1e3, 3)}
        if "gpps" in d:
            line[side]["gpps"] = round(d["gpps"], 3)
    line["failures"] = result["failures"]
    return line


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--output", default=None, help="write the full results as JSON here")
    ap.add_argument("--quality", type=int, default=90)
    ap.add_argument("--warmups", type=int, default=5)
    ap.add_argument("--repetitions", type=int, default=20)
    ap.add_argument("--tile-size", type=lambda v: [int(x) for x in v.split("x")],
                    default=[256, 128])
    ap.add_argument("--codec-module", default="fastvid")
    ap.add_argument("--samples", default=",".join(s.id for s in GATE_SAMPLES))
    args = ap.parse_args(argv)
    if args.repetitions < 20:
        ap.error("--repetitions must be at least 20")
    if args.warmups < 1:
        ap.error("--warmups must be positive")

    import torch
    codec = importlib.import_module(args.codec_module)
    wanted = set(args.samples.split(","))
    results, failures = [], []
    started = time.time()
    for sample in GATE_SAMPLES:
        if sample.id not in wanted:
            continue
        result = evaluate_sample(sample, codec, torch, args)
        result["failures"] = performance_failures(result)
        failures += [f"{sample.id}: {f}" for f in result["failures"]]
        results.append(result)
        print(json.dumps(summary_line(result)), flush=True)
    report = {"codec_module": args.codec_module, "device": torch.cuda.get_device_name(0),
              "quality": args.quality, "tile_size": args.tile_size, "data": "synthetic",
              "passed": not failures, "failures": failures,
              "elapsed_seconds": round(time.time() - started, 1), "results": results}
    if args.output:
        with open(args.output, "w") as fh:
            json.dump(report, fh, indent=1)
    print(json.dumps({"passed": report["passed"], "failures": failures,
                      "samples": [r["id"] for r in results]}))
    return 0 if not failures else 1


if __name__ == "__main__":
    sys.exit(main())
